// Multigrid set-up kernels and general sparse kernels on gfx950:
//  - Galerkin coarse operators R A P of 25-point stencils
//    (pymatting/preconditioner/vcycle.py:62-71, coarse_A = downsample.A.upsample),
//  - diagonal / float32 copies of stored stencils,
//  - the dense inverse used by the coarsest direct solve (vcycle.py:59-60),
//  - CSR SpMV for CG on arbitrary matrices (cg.py:68-71, 85-88).
#include <cstdarg>
#include <cstring>
#include <mutex>
#include <unordered_map>
#include <vector>

#include "common.h"
#include "engine.h"
#include "internal.h"

namespace hm {

// ------------------------------------------------------- kernel attributes

// Raises a kernel's dynamic LDS limit once per (kernel, device); threads that
// race here wait for the first one's call.
int set_max_dynamic_lds(const void* kernel, int bytes) {
    static std::mutex mu;
    static std::unordered_map<const void*, std::vector<int>> done;
    int dev = 0;
    HM_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lock(mu);
    std::vector<int>& devs = done[kernel];
    for (int d : devs)
        if (d == dev) return HM_OK;
    HM_HIP(hipFuncSetAttribute(kernel, hipFuncAttributeMaxDynamicSharedMemorySize, bytes));
    devs.push_back(dev);
    return HM_OK;
}

// ------------------------------------------------------------ device memory

// Device allocations go through a caching pool: a solve allocates dozens of
// buffers and hipMalloc/hipFree (which synchronises the device) would
// dominate small and medium solves. Freed blocks are kept per rounded size
// and reused in stream order; the library issues all work on the caller's
// stream, so a reused block's earlier uses are ordered before its new ones.
namespace {
std::mutex g_pool_mu;
std::unordered_map<size_t, std::vector<void*>> g_pool_free;
std::unordered_map<void*, size_t> g_pool_size;

size_t pool_round(size_t bytes) {
    if (bytes < 256) return 256;
    if (bytes < (1u << 20)) {
        size_t r = 256;
        while (r < bytes) r <<= 1;
        return r;
    }
    const size_t mb2 = size_t(2) << 20;
    return (bytes + mb2 - 1) / mb2 * mb2;
}
}  // namespace

int dev_alloc(void** p, size_t bytes) {
    *p = nullptr;
    const size_t r = pool_round(bytes);
    {
        std::lock_guard<std::mutex> lk(g_pool_mu);
        auto it = g_pool_free.find(r);
        if (it != g_pool_free.end() && !it->second.empty()) {
            *p = it->second.back();
            it->second.pop_back();
            return HM_OK;
        }
    }
    hipError_t e = hipMalloc(p, r);
    if (e != hipSuccess) {
        // Release cached blocks and retry once.
        pool_trim();
        e = hipMalloc(p, r);
    }
    if (e != hipSuccess) {
        set_error("hipMalloc(%zu) failed: %s", r, hipGetErrorString(e));
        *p = nullptr;
        return HM_ERR_HIP;
    }
    std::lock_guard<std::mutex> lk(g_pool_mu);
    g_pool_size[*p] = r;
    return HM_OK;
}

void dev_free(void* p) {
    if (!p) return;
    std::lock_guard<std::mutex> lk(g_pool_mu);
    auto it = g_pool_size.find(p);
    if (it == g_pool_size.end()) return;
    g_pool_free[it->second].push_back(p);
}

void pool_trim() {
    std::vector<void*> victims;
    {
        std::lock_guard<std::mutex> lk(g_pool_mu);
        for (auto& kv : g_pool_free) {
            for (void* q : kv.second) {
                victims.push_back(q);
                g_pool_size.erase(q);
            }
            kv.second.clear();
        }
    }
    if (!victims.empty()) (void)hipDeviceSynchronize();
    for (void* q : victims) (void)hipFree(q);
}

int make_reducer(Reducer* r, int max_blocks, int nvals) {
    r->max_blocks = max_blocks;
    r->upd_blocks = 0;
    HM_TRY(dev_alloc(reinterpret_cast<void**>(&r->partials),
                     sizeof(double) * (size_t)max_blocks * nvals));
    return HM_OK;
}

void free_reducer(Reducer* r) {
    dev_free(r->partials);
    r->partials = nullptr;
}

// ------------------------------------------------------------ Galerkin R A P

// Pairs (e_a, e_b) of fine offsets around coincident coarse points that keep
// the fine coupling b - a = 2 D + e_b - e_a inside the 5-point reach, for a
// coarse offset D in {-2..2}. Counts per D: 1, 6, 9, 6, 1.
__device__ __forceinline__ double w1(int e) { return e == 0 ? 0.5 : 0.25; }

// (R A P)_{I,J} for J = I + (Dy, Dx) with the bilinear prolongation P (weights
// w1(ey) w1(ex) from fine point 2I + e to coarse point I) and R = P^T: the sum
// over fine pairs a = 2I + ea, b = 2J + eb with b - a inside the 5 x 5 stencil,
// ea (then eb) ascending per dimension, rows outer. Compile-time offsets and
// clamped (always valid) addresses with out-of-grid pairs selected away let
// every gather of a point issue at once.
template <int Dy, int Dx>
__device__ __forceinline__ double galerkin_entry(const double* __restrict__ Af, long long nf,
                                                 int hf, int wf, int Y, int X) {
    double sum = 0.0;
#pragma unroll
    for (int eay = -1; eay <= 1; ++eay) {
#pragma unroll
        for (int eby = -1; eby <= 1; ++eby) {
            const int dyf = 2 * Dy + eby - eay;
            if (dyf < -2 || dyf > 2) continue;
            const int ay = 2 * Y + eay, by = 2 * (Y + Dy) + eby;
            const bool oky = ay >= 0 && ay < hf && by >= 0 && by < hf;
            const int ayc = ay < 0 ? 0 : (ay >= hf ? hf - 1 : ay);
            const double wy = w1(eay) * w1(eby);
#pragma unroll
            for (int eax = -1; eax <= 1; ++eax) {
#pragma unroll
                for (int ebx = -1; ebx <= 1; ++ebx) {
                    const int dxf = 2 * Dx + ebx - eax;
                    if (dxf < -2 || dxf > 2) continue;
                    const int ax = 2 * X + eax, bx = 2 * (X + Dx) + ebx;
                    const bool ok = oky && ax >= 0 && ax < wf && bx >= 0 && bx < wf;
                    const int axc = ax < 0 ? 0 : (ax >= wf ? wf - 1 : ax);
                    const double v =
                        sym_coef(Af, nf, wf, (long long)ayc * wf + axc, (dyf + 2) * 5 + dxf + 2);
                    sum += ok ? wy * (w1(eax) * w1(ebx)) * v : 0.0;
                }
            }
        }
    }
    return sum;
}

// One thread per (coarse point, stored coarse slot s >= 12); both operators
// use the symmetric storage of engine.h. The same store writes the float32
// planes, and slot 12 the diagonal and inverse diagonal (float64 and
// float32), with the operations of k_dia_finish.
__global__ void k_galerkin(const double* __restrict__ Af, int hf, int wf,
                           double* __restrict__ Ac, float* __restrict__ Ac32,
                           double* __restrict__ d64, float* __restrict__ d32,
                           double* __restrict__ id64, float* __restrict__ id32, int hc, int wc) {
    const long long nc = (long long)hc * wc, nf = (long long)hf * wf;
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= kSymSlots * nc) return;
    const int s = 12 + (int)(t / nc);
    const long long I = t - (long long)(s - 12) * nc;
    const int Y = (int)(I / wc), X = (int)(I - (long long)Y * wc);
    const int Dy = s / 5 - 2, Dx = s % 5 - 2;
    double sum = 0.0;
    if (Y + Dy >= 0 && Y + Dy < hc && X + Dx >= 0 && X + Dx < wc) {
        switch (s) {
#define HM_GAL(S_) \
    case S_: sum = galerkin_entry<S_ / 5 - 2, S_ % 5 - 2>(Af, nf, hf, wf, Y, X); break;
            HM_GAL(12) HM_GAL(13) HM_GAL(14) HM_GAL(15) HM_GAL(16) HM_GAL(17) HM_GAL(18)
            HM_GAL(19) HM_GAL(20) HM_GAL(21) HM_GAL(22) HM_GAL(23) HM_GAL(24)
#undef HM_GAL
            default: break;
        }
    }
    Ac[t] = sum;
    Ac32[t] = (float)sum;
    if (s == 12) {
        d64[I] = sum;
        d32[I] = (float)sum;
        id64[I] = 1.0 / sum;
        id32[I] = (float)(1.0 / sum);
    }
}

int dia_galerkin(const hm_dia* fine, hm_dia* coarse, hipStream_t s) {
    const long long total = kSymSlots * coarse->n;
    k_galerkin<<<div_up(total, 256), 256, 0, s>>>(fine->c64, fine->h, fine->w, coarse->c64,
                                                  coarse->c32, coarse->diag64, coarse->diag32,
                                                  coarse->idiag64, coarse->idiag32, coarse->h,
                                                  coarse->w);
    HM_LAUNCH_CHECK();
    return HM_OK;
}

__global__ void k_dia_finish(const double* __restrict__ c64, float* __restrict__ c32,
                             double* __restrict__ d64, float* __restrict__ d32,
                             double* __restrict__ id64, float* __restrict__ id32, long long n) {
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= kSymSlots * n) return;
    const double v = c64[t];
    c32[t] = (float)v;
    if (t < n) {  // stored slot 0 is the diagonal (s = 12)
        d64[t] = v;
        d32[t] = (float)v;
        id64[t] = 1.0 / v;
        id32[t] = (float)(1.0 / v);
    }
}

// Full [25][n] slot-major layout <-> symmetric [13][n] storage.
__global__ void k_dia_pack(const double* __restrict__ full, double* __restrict__ sym, long long n) {
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < kSymSlots * n) sym[t] = full[12 * n + t];
}

__global__ void k_dia_unpack(const double* __restrict__ sym, int h, int w,
                             double* __restrict__ full) {
    const long long n = (long long)h * w;
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 25 * n) return;
    const int s = (int)(t / n);
    const long long i = t - (long long)s * n;
    const int y = (int)(i / w), x = (int)(i - (long long)y * w);
    const int yy = y + s / 5 - 2, xx = x + s % 5 - 2;
    full[t] = (yy >= 0 && yy < h && xx >= 0 && xx < w) ? sym_coef(sym, n, w, i, s) : 0.0;
}

int dia_pack(const double* full, hm_dia* A, hipStream_t s) {
    k_dia_pack<<<div_up(kSymSlots * A->n, 256), 256, 0, s>>>(full, A->c64, A->n);
    HM_LAUNCH_CHECK();
    return HM_OK;
}

int dia_unpack(const hm_dia* A, double* full, hipStream_t s) {
    k_dia_unpack<<<div_up(25 * A->n, 256), 256, 0, s>>>(A->c64, A->h, A->w, full);
    HM_LAUNCH_CHECK();
    return HM_OK;
}

int dia_finish(hm_dia* A, hipStream_t s) {
    k_dia_finish<<<div_up(kSymSlots * A->n, 256), 256, 0, s>>>(
        A->c64, A->c32, A->diag64, A->diag32, A->idiag64, A->idiag32, A->n);
    HM_LAUNCH_CHECK();
    return HM_OK;
}

// ------------------------------------------------------- coarsest direct solve

// [A | I] (n x 2n) from a stored 25-point stencil or from a dense matrix.
__global__ void k_aug_from_dia(const double* __restrict__ coef, int h, int w,
                               double* __restrict__ M) {
    const int n = h * w, n2 = 2 * n;
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)n * n2) return;
    const int r = (int)(t / n2), c = (int)(t - (long long)r * n2);
    double v = 0.0;
    if (c < n) {
        const int ry = r / w, rx = r % w, cy = c / w, cx = c % w;
        const int dy = cy - ry, dx = cx - rx;
        if (dy >= -2 && dy <= 2 && dx >= -2 && dx <= 2)
            v = sym_coef(coef, (long long)n, w, (long long)r, (dy + 2) * 5 + dx + 2);
    } else {
        v = (c - n == r) ? 1.0 : 0.0;
    }
    M[t] = v;
}

__global__ void k_aug_from_dense(const double* __restrict__ A, int n, double* __restrict__ M) {
    const int n2 = 2 * n;
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)n * n2) return;
    const int r = (int)(t / n2), c = (int)(t - (long long)r * n2);
    M[t] = c < n ? A[(long long)r * n + c] : ((c - n == r) ? 1.0 : 0.0);
}

// Gauss-Jordan elimination with partial pivoting on [A | I] (global memory),
// one workgroup. Writes the inverse in float64 and float32.
__global__ __launch_bounds__(1024) void k_gauss_jordan(int n, double* __restrict__ M,
                                                       double* __restrict__ inv64,
                                                       float* __restrict__ inv32,
                                                       int* __restrict__ status) {
    const int n2 = 2 * n;
    __shared__ double fac[kMaxDirect];
    __shared__ double best_v[1024];
    __shared__ int best_i[1024];
    for (int k = 0; k < n; ++k) {
        double bv = -1.0;
        int bi = k;
        for (int r = k + threadIdx.x; r < n; r += blockDim.x) {
            const double a = fabs(M[(long long)r * n2 + k]);
            if (a > bv) { bv = a; bi = r; }
        }
        best_v[threadIdx.x] = bv;
        best_i[threadIdx.x] = bi;
        __syncthreads();
        // Tree argmax; ties keep the lower row like a serial scan.
        for (int off = blockDim.x / 2; off > 0; off >>= 1) {
            if ((int)threadIdx.x < off) {
                const double ov = best_v[threadIdx.x + off];
                const int oi = best_i[threadIdx.x + off];
                if (ov > best_v[threadIdx.x] || (ov == best_v[threadIdx.x] && oi < best_i[threadIdx.x])) {
                    best_v[threadIdx.x] = ov;
                    best_i[threadIdx.x] = oi;
                }
            }
            __syncthreads();
        }
        if (threadIdx.x == 0 && !(best_v[0] > 0.0)) *status = 1;
        __syncthreads();
        const int piv = best_i[0];
        if (piv != k)
            for (int c = threadIdx.x; c < n2; c += blockDim.x) {
                const double t0 = M[(long long)k * n2 + c];
                M[(long long)k * n2 + c] = M[(long long)piv * n2 + c];
                M[(long long)piv * n2 + c] = t0;
            }
        __syncthreads();
        const double inv_p = 1.0 / M[(long long)k * n2 + k];
        __syncthreads();
        for (int c = threadIdx.x; c < n2; c += blockDim.x) M[(long long)k * n2 + c] *= inv_p;
        for (int r = threadIdx.x; r < n; r += blockDim.x) fac[r] = M[(long long)r * n2 + k];
        __syncthreads();
        for (long long t = threadIdx.x; t < (long long)n * n2; t += blockDim.x) {
            const int r = (int)(t / n2), c = (int)(t - (long long)r * n2);
            if (r != k) M[t] -= fac[r] * M[(long long)k * n2 + c];
        }
        __syncthreads();
    }
    for (long long t = threadIdx.x; t < (long long)n * n; t += blockDim.x) {
        const int r = (int)(t / n), c = (int)(t - (long long)r * n);
        const double v = M[(long long)r * n2 + n + c];
        inv64[t] = v;
        inv32[t] = (float)v;
    }
}

// The same elimination for n <= kGjLds with [A | I] held in LDS: one wave
// finds the pivot (one row per lane, ties to the lower row), three barriers
// per column.
constexpr int kGjLds = 64;

__global__ __launch_bounds__(256) void k_gauss_jordan_lds(int n, const double* __restrict__ M,
                                                          double* __restrict__ inv64,
                                                          float* __restrict__ inv32,
                                                          int* __restrict__ status) {
    extern __shared__ double S[];  // n x 2n
    __shared__ double fac[kGjLds];
    __shared__ int piv_s;
    const int n2 = 2 * n, tid = threadIdx.x;
    for (int t = tid; t < n * n2; t += blockDim.x) S[t] = M[t];
    __syncthreads();
    for (int k = 0; k < n; ++k) {
        if (tid < 64) {
            const int r = k + tid;
            double v = r < n ? fabs(S[r * n2 + k]) : -1.0;
            if (!(v >= 0.0)) v = -1.0;  // NaN never pivots
            int idx = r;
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) {
                const double ov = __shfl_xor(v, off);
                const int oi = __shfl_xor(idx, off);
                if (ov > v || (ov == v && oi < idx)) {
                    v = ov;
                    idx = oi;
                }
            }
            if (tid == 0) {
                piv_s = idx;
                if (!(v > 0.0)) *status = 1;
            }
        }
        __syncthreads();
        const int piv = piv_s < n ? piv_s : k;
        if (piv != k)
            for (int c = tid; c < n2; c += blockDim.x) {
                const double t0 = S[k * n2 + c];
                S[k * n2 + c] = S[piv * n2 + c];
                S[piv * n2 + c] = t0;
            }
        __syncthreads();
        const double inv_p = 1.0 / S[k * n2 + k];
        if (tid < n) fac[tid] = S[tid * n2 + k];
        __syncthreads();
        for (int c = tid; c < n2; c += blockDim.x) S[k * n2 + c] *= inv_p;
        __syncthreads();
        for (int t = tid; t < n * n2; t += blockDim.x) {
            const int r = t / n2, c = t - r * n2;
            if (r != k) S[t] -= fac[r] * S[k * n2 + c];
        }
        __syncthreads();
    }
    for (int t = tid; t < n * n; t += blockDim.x) {
        const int r = t / n, c = t - r * n;
        const double v = S[r * n2 + n + c];
        inv64[t] = v;
        inv32[t] = (float)v;
    }
}

// status_out (a device int, nullable): the kernel's singularity flag is left
// there for the caller to read at its own next synchronisation; without it
// the flag is read back here.
static int gauss_jordan(int n, double* M, double* inv64, float* inv32, hipStream_t s,
                        int* status_out) {
    int* status = status_out;
    if (!status) HM_TRY(dev_alloc(reinterpret_cast<void**>(&status), sizeof(int)));
    HM_HIP(hipMemsetAsync(status, 0, sizeof(int), s));
    if (n <= kGjLds) {
        HM_TRY(set_max_dynamic_lds(reinterpret_cast<const void*>(&k_gauss_jordan_lds),
                                   (int)sizeof(double) * 2 * kGjLds * kGjLds));
        k_gauss_jordan_lds<<<1, 256, sizeof(double) * 2 * n * n, s>>>(n, M, inv64, inv32, status);
    } else {
        k_gauss_jordan<<<1, 1024, 0, s>>>(n, M, inv64, inv32, status);
    }
    HM_LAUNCH_CHECK();
    if (status_out) return HM_OK;
    int host_status = 0;
    HM_HIP(hipMemcpyAsync(&host_status, status, sizeof(int), hipMemcpyDeviceToHost, s));
    HM_HIP(hipStreamSynchronize(s));
    dev_free(status);
    if (host_status) {
        set_error("coarsest-grid matrix is singular");
        return HM_ERR_VALUE;
    }
    return HM_OK;
}

int dense_inverse(const hm_dia* A, double* inv64, float* inv32, hipStream_t s, int* status) {
    const int n = (int)A->n;
    if (n > kMaxDirect) {
        set_error("direct solve of a %d-point coarsest grid exceeds the %d-point limit", n,
                  kMaxDirect);
        return HM_ERR_UNSUPPORTED;
    }
    double* M = nullptr;
    HM_TRY(dev_alloc(reinterpret_cast<void**>(&M), sizeof(double) * (size_t)n * 2 * n));
    k_aug_from_dia<<<div_up((long long)n * 2 * n, 256), 256, 0, s>>>(A->c64, A->h, A->w, M);
    HM_LAUNCH_CHECK();
    int rc = gauss_jordan(n, M, inv64, inv32, s, status);
    dev_free(M);
    return rc;
}

int dense_inverse_from_dense(const double* A, int n, double* inv64, float* inv32, hipStream_t s) {
    if (n > kMaxDirect) {
        set_error("direct solve of a %d x %d matrix exceeds the %d limit", n, n, kMaxDirect);
        return HM_ERR_UNSUPPORTED;
    }
    double* M = nullptr;
    HM_TRY(dev_alloc(reinterpret_cast<void**>(&M), sizeof(double) * (size_t)n * 2 * n));
    k_aug_from_dense<<<div_up((long long)n * 2 * n, 256), 256, 0, s>>>(A, n, M);
    HM_LAUNCH_CHECK();
    int rc = gauss_jordan(n, M, inv64, inv32, s, nullptr);
    dev_free(M);
    return rc;
}

// ----------------------------------------------------------------------- CSR

__global__ __launch_bounds__(kBlock) void k_csr_apply(const int64_t* __restrict__ indptr,
                                                      const int64_t* __restrict__ indices,
                                                      const double* __restrict__ values,
                                                      long long n, const double* __restrict__ x,
                                                      double* __restrict__ y, CgState* st,
                                                      Reducer red, int kind) {
    if (solve_done(st)) return;
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    double acc = 0.0;
    if (i < n) {
        double s = 0.0;
        for (int64_t k = indptr[i]; k < indptr[i + 1]; ++k) s += values[k] * x[indices[k]];
        y[i] = s;
        acc = x[i] * s;
    }
    if (st != nullptr && red.partials != nullptr) {
        double v[1] = {acc};
        block_partials<kBlock, 1>(v, red);
    }
}

int csr_apply(const hm_csr* A, const double* x, double* y, CgState* st, Reducer* red, int kind,
              hipStream_t s) {
    const unsigned g = div_up(A->n, kBlock);
    Reducer r = red ? *red : Reducer{};
    if (red && (int)g > red->max_blocks) {
        set_error("reducer too small");
        return HM_ERR_ARG;
    }
    k_csr_apply<<<g, kBlock, 0, s>>>(A->indptr, A->indices, A->values, A->n, x, y, red ? st : nullptr,
                                     r, kind);
    HM_LAUNCH_CHECK();
    if (red && st) return launch_finalize(1, r, g, st, kind, st, s);
    return HM_OK;
}

int launch_finalize(int nvals, const Reducer& red, unsigned nblocks, CgState* st, int kind,
                    const CgState* guard, hipStream_t s) {
    if ((int)nblocks > red.max_blocks) {
        set_error("reduction buffer too small for %u blocks", nblocks);
        return HM_ERR_ARG;
    }
    if (kind == kDotUpdateRz && (int)red.upd_blocks > red.max_blocks) {
        set_error("reduction buffer too small for %u blocks", red.upd_blocks);
        return HM_ERR_ARG;
    }
    if (nvals == 1 && kind != kDotUpdateRz)
        k_finalize<1><<<1, kFinBlock, 0, s>>>(red, nblocks, st, kind, guard);
    else
        k_finalize<2><<<1, kFinBlock, 0, s>>>(red, nblocks, st, kind, guard);
    HM_LAUNCH_CHECK();
    return HM_OK;
}

}  // namespace hm
