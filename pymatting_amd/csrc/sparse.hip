// General CSR matrices on gfx950: the canonical sum A + diag(c) that
// pymatting/laplacian/laplacian.py:49-55 forms for an arbitrary Laplacian
// (scipy's L + diags(c), tocsr(), sum_duplicates()).
//
// For each row, entries of A with the same column are summed in their stored
// order (starting from 0.0), c_i is added to the diagonal's sum, sums equal
// to zero are dropped and the remaining columns are written in ascending
// order: the values and layout scipy's csr_plus_csr leaves after
// sum_duplicates() sorts it. One 64-lane workgroup handles one row, staged
// in LDS; a row may hold up to kMaxRow entries.
#include "common.h"
#include "internal.h"

namespace hm {

namespace {

constexpr int kRowLanes = 64;
constexpr int kMaxRow = 2048;

// mode 0: counts[i] = number of stored entries of row i (and the largest row
// length seen, for the caller's check, in *too_long if above kMaxRow).
// mode 1: fills out_idx/out_val from out_ptr.
__global__ __launch_bounds__(kRowLanes) void k_csr_diag_sum(
    long long n, const int64_t* __restrict__ ptr, const int64_t* __restrict__ idx,
    const double* __restrict__ val, const double* __restrict__ c, int mode,
    int64_t* __restrict__ counts, int* __restrict__ too_long, const int64_t* __restrict__ out_ptr,
    int64_t* __restrict__ out_idx, double* __restrict__ out_val) {
    __shared__ int64_t s_col[kMaxRow + 1];
    __shared__ double s_val[kMaxRow + 1];
    __shared__ double s_sum[kMaxRow + 1];
    __shared__ unsigned char s_keep[kMaxRow + 1];
    const long long i = blockIdx.x + (long long)blockIdx.y * gridDim.x;
    if (i >= n) return;
    const int64_t p0 = ptr[i], p1 = ptr[i + 1];
    const double ci = c ? c[i] : 0.0;
    const long long m_a = p1 - p0;
    if (m_a > kMaxRow) {
        if (threadIdx.x == 0) atomicOr(too_long, 1);
        return;
    }
    // A zero-valued diagonal entry appended when c_i != 0 stands for the
    // diagonal term scipy adds; summed after A's own diagonal entries it
    // leaves their sum unchanged (x + 0.0 == x for every non-zero x).
    const int m = (int)m_a + (ci != 0.0 ? 1 : 0);
    for (int e = threadIdx.x; e < m; e += kRowLanes) {
        if (e < m_a) {
            s_col[e] = idx[p0 + e];
            s_val[e] = val[p0 + e];
        } else {
            s_col[e] = i;
            s_val[e] = 0.0;
        }
    }
    __syncthreads();
    for (int e = threadIdx.x; e < m; e += kRowLanes) {
        const int64_t col = s_col[e];
        bool first = true;
        for (int q = 0; q < e && first; ++q) first = s_col[q] != col;
        double sum = 0.0;
        if (first) {
            for (int q = e; q < m; ++q)
                if (s_col[q] == col) sum += s_val[q];
            if (col == i) sum += ci;
        }
        s_sum[e] = sum;
        s_keep[e] = first && sum != 0.0;
    }
    __syncthreads();
    if (mode == 0) {
        int cnt = 0;
        for (int e = threadIdx.x; e < m; e += kRowLanes) cnt += s_keep[e];
        for (int off = kRowLanes / 2; off > 0; off >>= 1) cnt += __shfl_xor(cnt, off, kRowLanes);
        if (threadIdx.x == 0) counts[i] = cnt;
        return;
    }
    const int64_t base = out_ptr[i];
    for (int e = threadIdx.x; e < m; e += kRowLanes) {
        if (!s_keep[e]) continue;
        const int64_t col = s_col[e];
        int rank = 0;
        for (int q = 0; q < m; ++q) rank += (s_keep[q] && s_col[q] < col) ? 1 : 0;
        out_idx[base + rank] = col;
        out_val[base + rank] = s_sum[e];
    }
}

// ------------------------------------------- reduced system (hard constraints)
//
// estimate_alpha_cf.py:57-70 on a CSR matrix L: L_U = L[U, :][:, U] and
// R = L[U, :][:, K] as scipy's fancy indexing leaves them (the rows of the
// unknown pixels in order; in each, the entries whose column is unknown, in
// stored order, renumbered by rank among the unknown pixels, duplicates and
// explicit zeros kept), and rhs = -(R m) with m = is_fg[K] (known columns of
// class 1), summed in stored order from 0.0 like scipy's CSR product.
// cls: 0 unknown, 1 foreground, 2 background.
__global__ void k_unknown_flags(const uint8_t* __restrict__ cls, long long n,
                                int64_t* __restrict__ flags) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) flags[i] = cls[i] == 0 ? 1 : 0;
}

__global__ void k_reduced_counts(long long n, const int64_t* __restrict__ ptr,
                                 const int64_t* __restrict__ idx,
                                 const uint8_t* __restrict__ cls,
                                 const int64_t* __restrict__ rank, int64_t* __restrict__ counts) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || cls[i] != 0) return;
    int64_t cnt = 0;
    for (int64_t p = ptr[i]; p < ptr[i + 1]; ++p) cnt += cls[idx[p]] == 0 ? 1 : 0;
    counts[rank[i]] = cnt;
}

__global__ void k_reduced_fill(long long n, const int64_t* __restrict__ ptr,
                               const int64_t* __restrict__ idx, const double* __restrict__ val,
                               const uint8_t* __restrict__ cls, const int64_t* __restrict__ rank,
                               const int64_t* __restrict__ ptr_u, int64_t* __restrict__ idx_u,
                               double* __restrict__ val_u, double* __restrict__ rhs_u) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || cls[i] != 0) return;
    const long long row = rank[i];
    int64_t q = ptr_u[row];
    double sum = 0.0;
    for (int64_t p = ptr[i]; p < ptr[i + 1]; ++p) {
        const int64_t j = idx[p];
        const uint8_t cj = cls[j];
        if (cj == 0) {
            idx_u[q] = rank[j];
            val_u[q] = val[p];
            ++q;
        } else {
            sum += val[p] * (cj == 1 ? 1.0 : 0.0);
        }
    }
    rhs_u[row] = -sum;
}

// alpha = clip(x, 0, 1) with x = trimap on known pixels and x_U on unknown
// ones (estimate_alpha_cf.py:72-76).
__global__ void k_reduced_finalize(const uint8_t* __restrict__ cls, long long n,
                                   const int64_t* __restrict__ rank, const double* __restrict__ xu,
                                   const double* __restrict__ trimap, double* __restrict__ alpha) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double v = cls[i] == 0 ? xu[rank[i]] : trimap[i];
    v = v < 0.0 ? 0.0 : (v > 1.0 ? 1.0 : v);
    alpha[i] = v;
}

struct DevTmp {
    void* p = nullptr;
    ~DevTmp() { dev_free(p); }
};

// rank[i] = number of unknown pixels before i (n + 1 entries; the last, also
// returned in *n_u, counts them all).
int unknown_rank(const uint8_t* cls, long long n, int64_t* rank, int64_t* n_u, hipStream_t s) {
    DevTmp flags;
    HM_TRY(dev_alloc(&flags.p, sizeof(int64_t) * (n > 0 ? n : 1)));
    if (n > 0) {
        k_unknown_flags<<<div_up(n, 256), 256, 0, s>>>(cls, n, static_cast<int64_t*>(flags.p));
        HM_LAUNCH_CHECK();
    }
    return exclusive_offsets(static_cast<int64_t*>(flags.p), rank, n, n_u, s);
}

inline dim3 row_grid(long long n) {
    const long long gx = n < 65535 ? (n > 0 ? n : 1) : 65535;
    return dim3((unsigned)gx, (unsigned)div_up(n > 0 ? n : 1, gx));
}

}  // namespace

}  // namespace hm

extern "C" {

int hm_csr_add_diagonal_indptr(int64_t n, const int64_t* indptr, const int64_t* indices,
                               const double* values, const double* diag, int64_t* out_indptr,
                               int64_t* nnz, void* stream) {
    using namespace hm;
    if (n < 0 || !indptr || !out_indptr || !nnz || (n > 0 && (!indices || !values))) {
        set_error("hm_csr_add_diagonal_indptr: invalid argument");
        return HM_ERR_ARG;
    }
    hipStream_t s = static_cast<hipStream_t>(stream);
    void* counts = nullptr;
    void* flag = nullptr;
    HM_TRY(dev_alloc(&counts, sizeof(int64_t) * (n > 0 ? n : 1)));
    if (dev_alloc(&flag, sizeof(int)) != HM_OK) {
        dev_free(counts);
        return HM_ERR_HIP;
    }
    int rc = HM_OK;
    int too_long = 0;
    hipError_t e = hipMemsetAsync(flag, 0, sizeof(int), s);
    if (e == hipSuccess && n > 0) {
        k_csr_diag_sum<<<row_grid(n), kRowLanes, 0, s>>>(
            n, indptr, indices, values, diag, 0, static_cast<int64_t*>(counts),
            static_cast<int*>(flag), nullptr, nullptr, nullptr);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(&too_long, flag, sizeof(int), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) {
        set_error("hm_csr_add_diagonal_indptr: %s", hipGetErrorString(e));
        rc = HM_ERR_HIP;
    } else if (too_long) {
        set_error("hm_csr_add_diagonal_indptr: a row holds more than %d entries", kMaxRow);
        rc = HM_ERR_UNSUPPORTED;
    } else {
        rc = exclusive_offsets(static_cast<int64_t*>(counts), out_indptr, n, nnz, s);
    }
    dev_free(counts);
    dev_free(flag);
    return rc;
}

int hm_csr_add_diagonal_fill(int64_t n, const int64_t* indptr, const int64_t* indices,
                             const double* values, const double* diag,
                             const int64_t* out_indptr, int64_t* out_indices, double* out_values,
                             void* stream) {
    using namespace hm;
    if (n < 0 || !indptr || !out_indptr || (n > 0 && (!indices || !values)) ||
        (n > 0 && (!out_indices || !out_values))) {
        set_error("hm_csr_add_diagonal_fill: invalid argument");
        return HM_ERR_ARG;
    }
    if (n == 0) return HM_OK;
    void* flag = nullptr;
    HM_TRY(dev_alloc(&flag, sizeof(int)));
    hipStream_t s = static_cast<hipStream_t>(stream);
    hipError_t e = hipMemsetAsync(flag, 0, sizeof(int), s);
    if (e == hipSuccess) {
        k_csr_diag_sum<<<row_grid(n), kRowLanes, 0, s>>>(n, indptr, indices, values, diag, 1,
                                                         nullptr, static_cast<int*>(flag),
                                                         out_indptr, out_indices, out_values);
        e = hipGetLastError();
    }
    dev_free(flag);
    if (e != hipSuccess) {
        set_error("hm_csr_add_diagonal_fill: %s", hipGetErrorString(e));
        return HM_ERR_HIP;
    }
    return HM_OK;
}

int hm_csr_reduced_indptr(int64_t n, const int64_t* indptr, const int64_t* indices,
                          const uint8_t* cls, int64_t* indptr_u, int64_t* n_u, int64_t* nnz_u,
                          void* stream) {
    using namespace hm;
    if (n < 1 || !indptr || !indices || !cls || !indptr_u || !n_u || !nnz_u) {
        set_error("hm_csr_reduced_indptr: invalid argument");
        return HM_ERR_ARG;
    }
    hipStream_t s = static_cast<hipStream_t>(stream);
    DevTmp rank, counts;
    HM_TRY(dev_alloc(&rank.p, sizeof(int64_t) * (n + 1)));
    HM_TRY(unknown_rank(cls, n, static_cast<int64_t*>(rank.p), n_u, s));
    HM_TRY(dev_alloc(&counts.p, sizeof(int64_t) * (*n_u > 0 ? *n_u : 1)));
    k_reduced_counts<<<div_up(n, 256), 256, 0, s>>>(n, indptr, indices, cls,
                                                     static_cast<int64_t*>(rank.p),
                                                     static_cast<int64_t*>(counts.p));
    HM_LAUNCH_CHECK();
    return exclusive_offsets(static_cast<int64_t*>(counts.p), indptr_u, *n_u, nnz_u, s);
}

int hm_csr_reduced_fill(int64_t n, const int64_t* indptr, const int64_t* indices,
                        const double* values, const uint8_t* cls, const int64_t* indptr_u,
                        int64_t* indices_u, double* values_u, double* rhs_u, void* stream) {
    using namespace hm;
    if (n < 1 || !indptr || !indices || !values || !cls || !indptr_u || !indices_u ||
        !values_u || !rhs_u) {
        set_error("hm_csr_reduced_fill: invalid argument");
        return HM_ERR_ARG;
    }
    hipStream_t s = static_cast<hipStream_t>(stream);
    DevTmp rank;
    int64_t n_u = 0;
    HM_TRY(dev_alloc(&rank.p, sizeof(int64_t) * (n + 1)));
    HM_TRY(unknown_rank(cls, n, static_cast<int64_t*>(rank.p), &n_u, s));
    k_reduced_fill<<<div_up(n, 256), 256, 0, s>>>(n, indptr, indices, values, cls,
                                                   static_cast<int64_t*>(rank.p), indptr_u,
                                                   indices_u, values_u, rhs_u);
    HM_LAUNCH_CHECK();
    return HM_OK;
}

int hm_reduced_finalize(int64_t n, const uint8_t* cls, const double* x_u, const double* trimap,
                        double* alpha, void* stream) {
    using namespace hm;
    if (n < 1 || !cls || !x_u || !trimap || !alpha) {
        set_error("hm_reduced_finalize: invalid argument");
        return HM_ERR_ARG;
    }
    hipStream_t s = static_cast<hipStream_t>(stream);
    DevTmp rank;
    int64_t n_u = 0;
    HM_TRY(dev_alloc(&rank.p, sizeof(int64_t) * (n + 1)));
    HM_TRY(unknown_rank(cls, n, static_cast<int64_t*>(rank.p), &n_u, s));
    k_reduced_finalize<<<div_up(n, 256), 256, 0, s>>>(cls, n, static_cast<int64_t*>(rank.p), x_u,
                                                       trimap, alpha);
    HM_LAUNCH_CHECK();
    return HM_OK;
}

}  // extern "C"
