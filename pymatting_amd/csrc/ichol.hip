// Thresholded incomplete Cholesky preconditioner
// (pymatting/preconditioner/ichol.py:6-181).
//
// The factorisation is sequential by construction (column j needs every
// earlier column that reaches row j), so it runs on the host, in one thread,
// as the reference's does; its device form (the factor and the solves' level
// schedules) is uploaded on the first application. Applying the factor, z = L^-T L^-1 r, is what PCG
// repeats every iteration: it runs on the GPU as two level-scheduled
// triangular solves. Row i of a solve belongs to level 1 + max(level of the
// rows it reads); one 1024-thread workgroup sweeps the levels in order, the
// rows of a level in parallel, with a barrier between levels. The solution
// vector lives in LDS when it fits (up to 16384 rows), else in global memory.
//
// Both solves sum a row's terms in ascending column order with separate
// multiplies and subtractions (compiled with -ffp-contract=off), which is the
// operation order of the reference's column-oriented forward substitution
// (ichol.py:116-129) and row-oriented back substitution (:132-145), so the
// results are bit-identical to them.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "common.h"
#include "internal.h"

struct hm_ichol {
    long long n = 0, nnz = 0;
    // Host copy of the factor in the reference's CSC layout (column j: the
    // diagonal first, then the rows below it in ascending order).
    std::vector<double> Lv;
    std::vector<int64_t> Lr, Lp;
    // Device forms. Forward: row i of L without its diagonal (columns
    // ascending). Backward: column i of L without its diagonal (rows
    // ascending), i.e. row i of L^T. Rows of each solve are listed level by
    // level.
    int* f_ptr = nullptr;
    int* f_idx = nullptr;
    double* f_val = nullptr;
    int* f_lev = nullptr;
    int* f_rows = nullptr;
    int f_nlev = 0;
    int* b_ptr = nullptr;
    int* b_idx = nullptr;
    double* b_val = nullptr;
    int* b_lev = nullptr;
    int* b_rows = nullptr;
    int b_nlev = 0;
    double* diag = nullptr;
    double* y = nullptr;  // forward solution, input of the backward solve
    // Host forms of the above, made with the factor and uploaded (then
    // released) by the first application.
    std::vector<int> hf_ptr, hf_idx, hf_lev, hf_rows, hb_ptr, hb_idx, hb_lev, hb_rows;
    std::vector<double> hf_val, hb_val, h_diag;
    bool uploaded = false;
};

namespace hm {

namespace {

// Left-looking factorisation of the lower triangle of a CSC matrix (columns
// ascending rows) into Lv/Lr/Lp with the reference's arithmetic: the
// diagonal gets `shift` added, column j accumulates A's entries below the
// diagonal and subtracts L_ik L_jk for every earlier column k with an entry
// in row j, visited in the order of a per-row list into which column k is
// pushed (at the front) when its next unvisited entry lies in that row. Its
// entries are then divided by the pivot, visited in ascending row order; the
// square of each one updates the diagonal of its row (before or after the
// discard test, per keep_discarded), and only those above both thresholds
// are stored. Returns nnz, -1 for a pivot <= 0, -2 when more than max_nnz
// entries would be stored.
long long factor(long long n, const int64_t* Ap, const int64_t* Ar, const double* Av,
                 double discard, double shift, long long max_nnz, double rel_discard,
                 bool keep_discarded, std::vector<double>& Lv, std::vector<int64_t>& Lr,
                 std::vector<int64_t>& Lp) {
    std::vector<double> pivot(n, shift), col_abs(n, 0.0), work(n, 0.0);
    std::vector<int64_t> sub_start(n, 0);   // first entry below A's diagonal in column j
    std::vector<int64_t> cursor(n, 0);      // next unvisited entry of column k of L
    std::vector<int64_t> list_head(n, -1);  // columns whose next entry lies in row i
    std::vector<int64_t> list_next(n, -1);
    std::vector<int64_t> touched;
    std::vector<char> mark(n, 0);
    touched.reserve(64);
    for (long long j = 0; j < n; ++j)
        for (int64_t p = Ap[j]; p < Ap[j + 1]; ++p) {
            const int64_t i = Ar[p];
            if (i == j) {
                pivot[j] += Av[p];
                sub_start[j] = p + 1;
            }
            if (i >= j) col_abs[j] += std::fabs(Av[p]);
        }
    Lv.clear();
    Lr.clear();
    Lp.assign(n + 1, 0);
    auto touch = [&](int64_t i) {
        if (!mark[i]) {
            mark[i] = 1;
            touched.push_back(i);
        }
    };
    for (long long j = 0; j < n; ++j) {
        for (int64_t p = sub_start[j]; p < Ap[j + 1]; ++p) {
            const int64_t i = Ar[p];
            if (Av[p] != 0.0 && i > j) {
                work[i] += Av[p];
                touch(i);
            }
        }
        for (int64_t k = list_head[j]; k != -1;) {
            const int64_t after = list_next[k];
            int64_t c = cursor[k];
            const int64_t end = Lp[k + 1];
            const double Ljk = Lv[c++];
            if (c < end) {
                cursor[k] = c;
                const int64_t row = Lr[c];
                list_next[k] = list_head[row];
                list_head[row] = k;
                for (int64_t q = c; q < end; ++q) {
                    work[Lr[q]] -= Lv[q] * Ljk;
                    touch(Lr[q]);
                }
            }
            k = after;
        }
        if (pivot[j] <= 0.0) return -1;
        if ((long long)Lv.size() + 1 + (long long)touched.size() > max_nnz) return -2;
        const double d = std::sqrt(pivot[j]);
        pivot[j] = d;
        Lv.push_back(d);
        Lr.push_back(j);
        cursor[j] = (int64_t)Lv.size();
        std::sort(touched.begin(), touched.end());
        const double rel = rel_discard * col_abs[j];
        for (const int64_t i : touched) {
            const double Lij = work[i] / d;
            if (keep_discarded) pivot[i] -= Lij * Lij;
            if (std::fabs(Lij) > discard && std::fabs(work[i]) > rel) {
                if (!keep_discarded) pivot[i] -= Lij * Lij;
                Lv.push_back(Lij);
                Lr.push_back(i);
            }
            work[i] = 0.0;
            mark[i] = 0;
        }
        touched.clear();
        Lp[j + 1] = (int64_t)Lv.size();
        if (Lp[j] + 1 < Lp[j + 1]) {
            const int64_t row = Lr[Lp[j] + 1];
            list_next[j] = list_head[row];
            list_head[row] = j;
        }
    }
    return (long long)Lv.size();
}

// Level schedule of a triangular solve whose row i reads the rows idx[ptr[i]
// .. ptr[i+1]); `forward` rows read only earlier rows, backward rows only
// later ones. Returns the level offsets and the rows level by level.
void schedule(long long n, const std::vector<int>& ptr, const std::vector<int>& idx, bool forward,
              std::vector<int>& lev_ptr, std::vector<int>& rows) {
    std::vector<int> level(n, 0);
    int nlev = 0;
    for (long long t = 0; t < n; ++t) {
        const long long i = forward ? t : n - 1 - t;
        int l = 0;
        for (int q = ptr[i]; q < ptr[i + 1]; ++q) l = std::max(l, level[idx[q]] + 1);
        level[i] = l;
        nlev = std::max(nlev, l + 1);
    }
    lev_ptr.assign(nlev + 1, 0);
    for (long long i = 0; i < n; ++i) ++lev_ptr[level[i] + 1];
    for (int l = 0; l < nlev; ++l) lev_ptr[l + 1] += lev_ptr[l];
    rows.assign(n, 0);
    std::vector<int> fill(lev_ptr.begin(), lev_ptr.end() - 1);
    for (long long i = 0; i < n; ++i) rows[fill[level[i]]++] = (int)i;
}

constexpr int kTrsvBlock = 1024;
constexpr int kTrsvLdsRows = 16384;  // solution vectors up to this size stay in LDS

// out_i = (rhs_i - sum_q val[q] out[idx[q]]) / diag[i] for the rows of each
// level in turn (terms in the stored order: ascending column). kLds keeps
// `out` in LDS and writes it back at the end.
template <bool kLds>
__global__ __launch_bounds__(kTrsvBlock) void k_trsv(int nlev, const int* __restrict__ lev,
                                                     const int* __restrict__ rows,
                                                     const int* __restrict__ ptr,
                                                     const int* __restrict__ idx,
                                                     const double* __restrict__ val,
                                                     const double* __restrict__ diag,
                                                     const double* __restrict__ rhs,
                                                     double* __restrict__ out, int n,
                                                     const CgState* guard) {
    if (solve_done(guard)) return;
    extern __shared__ double s_out[];
    double* v = kLds ? s_out : out;
    for (int l = 0; l < nlev; ++l) {
        for (int t = lev[l] + (int)threadIdx.x; t < lev[l + 1]; t += kTrsvBlock) {
            const int i = rows[t];
            double s = rhs[i];
            for (int q = ptr[i]; q < ptr[i + 1]; ++q) s -= val[q] * v[idx[q]];
            v[i] = s / diag[i];
        }
        __syncthreads();
    }
    if (kLds)
        for (int i = threadIdx.x; i < n; i += kTrsvBlock) out[i] = s_out[i];
}

template <typename T>
int upload(T** dst, const T* src, size_t count, hipStream_t s) {
    HM_TRY(dev_alloc(reinterpret_cast<void**>(dst), sizeof(T) * (count ? count : 1)));
    if (count) HM_HIP(hipMemcpyAsync(*dst, src, sizeof(T) * count, hipMemcpyHostToDevice, s));
    return HM_OK;
}

void release(hm_ichol* h) {
    if (!h) return;
    for (void* p : {(void*)h->f_ptr, (void*)h->f_idx, (void*)h->f_val, (void*)h->f_lev,
                    (void*)h->f_rows, (void*)h->b_ptr, (void*)h->b_idx, (void*)h->b_val,
                    (void*)h->b_lev, (void*)h->b_rows, (void*)h->diag, (void*)h->y})
        dev_free(p);
    delete h;
}

// Solve forms and level schedules of the host factor held in h (host
// vectors; upload() copies them to the device).
int build_host(hm_ichol* h) {
    const long long n = h->n;
    const std::vector<int64_t>& Lp = h->Lp;
    const std::vector<int64_t>& Lr = h->Lr;
    const std::vector<double>& Lv = h->Lv;
    std::vector<double>& diag = h->h_diag;
    diag.assign(n, 0.0);
    // Backward solve: the columns of L below the diagonal as they are.
    std::vector<int>& bptr = h->hb_ptr;
    std::vector<int>& bidx = h->hb_idx;
    std::vector<double>& bval = h->hb_val;
    bptr.assign(n + 1, 0);
    bidx.clear();
    bval.clear();
    bidx.reserve(Lv.size());
    bval.reserve(Lv.size());
    for (long long j = 0; j < n; ++j) {
        const int64_t p0 = Lp[j], p1 = Lp[j + 1];
        if (p1 <= p0 || Lr[p0] != j) {
            set_error("ichol factor: column %lld does not start with its diagonal", j);
            return HM_ERR_VALUE;
        }
        diag[j] = Lv[p0];
        for (int64_t p = p0 + 1; p < p1; ++p) {
            if (Lr[p] <= j || Lr[p] >= n || (p > p0 + 1 && Lr[p] <= Lr[p - 1])) {
                set_error("ichol factor: column %lld is not strictly lower triangular with "
                          "ascending rows", j);
                return HM_ERR_VALUE;
            }
            bidx.push_back((int)Lr[p]);
            bval.push_back(Lv[p]);
        }
        bptr[j + 1] = (int)bidx.size();
    }
    // Forward solve: the rows of L below the diagonal, columns ascending
    // (a transpose of the column form, filled in column order).
    std::vector<int>& fptr = h->hf_ptr;
    std::vector<int>& fidx = h->hf_idx;
    std::vector<double>& fval = h->hf_val;
    fptr.assign(n + 1, 0);
    fidx.assign(bidx.size(), 0);
    fval.assign(bidx.size(), 0.0);
    for (int q : bidx) ++fptr[q + 1];
    for (long long i = 0; i < n; ++i) fptr[i + 1] += fptr[i];
    {
        std::vector<int> fill(fptr.begin(), fptr.end() - 1);
        for (long long j = 0; j < n; ++j)
            for (int q = bptr[j]; q < bptr[j + 1]; ++q) {
                const int i = bidx[q];
                fidx[fill[i]] = (int)j;
                fval[fill[i]++] = bval[q];
            }
    }
    schedule(n, fptr, fidx, true, h->hf_lev, h->hf_rows);
    schedule(n, bptr, bidx, false, h->hb_lev, h->hb_rows);
    h->f_nlev = (int)h->hf_lev.size() - 1;
    h->b_nlev = (int)h->hb_lev.size() - 1;
    return HM_OK;
}

// Uploads the solve forms on s (the copies complete before this returns)
// and releases the host staging vectors.
int upload_device(hm_ichol* h, hipStream_t s) {
    const long long n = h->n;
    HM_TRY(upload(&h->f_ptr, h->hf_ptr.data(), h->hf_ptr.size(), s));
    HM_TRY(upload(&h->f_idx, h->hf_idx.data(), h->hf_idx.size(), s));
    HM_TRY(upload(&h->f_val, h->hf_val.data(), h->hf_val.size(), s));
    HM_TRY(upload(&h->f_lev, h->hf_lev.data(), h->hf_lev.size(), s));
    HM_TRY(upload(&h->f_rows, h->hf_rows.data(), h->hf_rows.size(), s));
    HM_TRY(upload(&h->b_ptr, h->hb_ptr.data(), h->hb_ptr.size(), s));
    HM_TRY(upload(&h->b_idx, h->hb_idx.data(), h->hb_idx.size(), s));
    HM_TRY(upload(&h->b_val, h->hb_val.data(), h->hb_val.size(), s));
    HM_TRY(upload(&h->b_lev, h->hb_lev.data(), h->hb_lev.size(), s));
    HM_TRY(upload(&h->b_rows, h->hb_rows.data(), h->hb_rows.size(), s));
    HM_TRY(upload(&h->diag, h->h_diag.data(), h->h_diag.size(), s));
    HM_TRY(dev_alloc(reinterpret_cast<void**>(&h->y), sizeof(double) * (n ? n : 1)));
    HM_HIP(hipStreamSynchronize(s));
    for (std::vector<int>* v : {&h->hf_ptr, &h->hf_idx, &h->hf_lev, &h->hf_rows, &h->hb_ptr,
                                &h->hb_idx, &h->hb_lev, &h->hb_rows})
        std::vector<int>().swap(*v);
    std::vector<double>().swap(h->hf_val);
    std::vector<double>().swap(h->hb_val);
    std::vector<double>().swap(h->h_diag);
    h->uploaded = true;
    return HM_OK;
}

}  // namespace

long long ichol_rows(const hm_ichol* h) { return h ? h->n : -1; }

// z = L^-T L^-1 r on the device (float64 vectors of h->n values). The
// device form is built on the first call.
int ichol_apply(const hm_ichol* hc, const double* r, double* z, const CgState* guard,
                hipStream_t s) {
    hm_ichol* h = const_cast<hm_ichol*>(hc);
    const int n = (int)h->n;
    if (n == 0) return HM_OK;
    if (!h->uploaded) HM_TRY(upload_device(h, s));
    if (n <= kTrsvLdsRows) {
        const size_t lds = sizeof(double) * n;
        HM_TRY(set_max_dynamic_lds(reinterpret_cast<const void*>(&k_trsv<true>),
                                   (int)(sizeof(double) * kTrsvLdsRows)));
        k_trsv<true><<<1, kTrsvBlock, lds, s>>>(h->f_nlev, h->f_lev, h->f_rows, h->f_ptr,
                                                h->f_idx, h->f_val, h->diag, r, h->y, n, guard);
        k_trsv<true><<<1, kTrsvBlock, lds, s>>>(h->b_nlev, h->b_lev, h->b_rows, h->b_ptr,
                                                h->b_idx, h->b_val, h->diag, h->y, z, n, guard);
    } else {
        k_trsv<false><<<1, kTrsvBlock, 0, s>>>(h->f_nlev, h->f_lev, h->f_rows, h->f_ptr,
                                               h->f_idx, h->f_val, h->diag, r, h->y, n, guard);
        k_trsv<false><<<1, kTrsvBlock, 0, s>>>(h->b_nlev, h->b_lev, h->b_rows, h->b_ptr,
                                               h->b_idx, h->b_val, h->diag, h->y, z, n, guard);
    }
    HM_LAUNCH_CHECK();
    return HM_OK;
}

}  // namespace hm

extern "C" {

int hm_ichol_factor(int64_t n, const int64_t* indptr, const int64_t* indices,
                    const double* values, double discard_threshold, double shift,
                    int64_t max_nnz, double relative_discard_threshold, int diag_keep_discarded,
                    int64_t* result, hm_ichol** out) {
    using namespace hm;
    if (!out || !result || n < 0 || (n > 0 && (!indptr || !indices || !values))) {
        set_error("hm_ichol_factor: invalid argument");
        return HM_ERR_ARG;
    }
    *out = nullptr;
    if (n >= (int64_t(1) << 31)) {
        set_error("hm_ichol_factor: %lld rows exceed the 32-bit row indices of the solves",
                  (long long)n);
        return HM_ERR_UNSUPPORTED;
    }
    for (int64_t j = 0; j < n; ++j)
        if (indptr[j + 1] < indptr[j]) {
            set_error("hm_ichol_factor: indptr is not non-decreasing");
            return HM_ERR_ARG;
        }
    for (int64_t p = 0; p < indptr[n]; ++p)
        if (indices[p] < 0 || indices[p] >= n) {
            set_error("hm_ichol_factor: row index out of range");
            return HM_ERR_ARG;
        }
    hm_ichol* h = new hm_ichol;
    h->n = n;
    const long long nnz = factor(n, indptr, indices, values, discard_threshold, shift, max_nnz,
                                 relative_discard_threshold, diag_keep_discarded != 0, h->Lv,
                                 h->Lr, h->Lp);
    *result = nnz;
    if (nnz < 0) {
        release(h);
        return HM_OK;
    }
    if (nnz >= (int64_t(1) << 31)) {
        release(h);
        set_error("hm_ichol_factor: %lld factor entries exceed 32-bit indices", nnz);
        return HM_ERR_UNSUPPORTED;
    }
    h->nnz = nnz;
    const int rc = build_host(h);
    if (rc != HM_OK) {
        release(h);
        return rc;
    }
    *out = h;
    return HM_OK;
}

int hm_ichol_create(int64_t n, const double* Lv, const int64_t* Lr, const int64_t* Lp,
                    hm_ichol** out) {
    using namespace hm;
    if (!out || n < 0 || !Lp || (n > 0 && Lp[n] > 0 && (!Lv || !Lr))) {
        set_error("hm_ichol_create: invalid argument");
        return HM_ERR_ARG;
    }
    *out = nullptr;
    if (n >= (int64_t(1) << 31) || Lp[n] >= (int64_t(1) << 31)) {
        set_error("hm_ichol_create: factor too large for 32-bit indices");
        return HM_ERR_UNSUPPORTED;
    }
    hm_ichol* h = new hm_ichol;
    h->n = n;
    h->nnz = Lp[n];
    h->Lp.assign(Lp, Lp + n + 1);
    h->Lr.assign(Lr, Lr + h->nnz);
    h->Lv.assign(Lv, Lv + h->nnz);
    for (int64_t j = 0; j < n; ++j)
        if (Lp[j + 1] < Lp[j] + 1 || Lr[Lp[j]] != j) {
            release(h);
            set_error("hm_ichol_create: column %lld does not start with its diagonal",
                      (long long)j);
            return HM_ERR_VALUE;
        }
    const int rc = build_host(h);
    if (rc != HM_OK) {
        release(h);
        return rc;
    }
    *out = h;
    return HM_OK;
}

int64_t hm_ichol_nnz(const hm_ichol* h) { return h ? h->nnz : -1; }

int hm_ichol_levels(const hm_ichol* h, int* forward, int* backward) {
    if (!h || !forward || !backward) {
        hm::set_error("hm_ichol_levels: null argument");
        return HM_ERR_ARG;
    }
    *forward = h->f_nlev;
    *backward = h->b_nlev;
    return HM_OK;
}

int hm_ichol_copy_factor(const hm_ichol* h, double* Lv, int64_t* Lr, int64_t* Lp) {
    if (!h || !Lp || (h->nnz > 0 && (!Lv || !Lr))) {
        hm::set_error("hm_ichol_copy_factor: null argument");
        return HM_ERR_ARG;
    }
    std::memcpy(Lp, h->Lp.data(), sizeof(int64_t) * (h->n + 1));
    if (h->nnz > 0) {
        std::memcpy(Lr, h->Lr.data(), sizeof(int64_t) * h->nnz);
        std::memcpy(Lv, h->Lv.data(), sizeof(double) * h->nnz);
    }
    return HM_OK;
}

int hm_ichol_apply(hm_ichol* h, const double* r, double* z, void* stream) {
    if (!h || (h->n > 0 && (!r || !z))) {
        hm::set_error("hm_ichol_apply: null argument");
        return HM_ERR_ARG;
    }
    return hm::ichol_apply(h, r, z, nullptr, static_cast<hipStream_t>(stream));
}

int hm_ichol_destroy(hm_ichol* h) {
    hm::release(h);
    return HM_OK;
}

}  // extern "C"
