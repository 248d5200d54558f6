// Thresholded incomplete Cholesky preconditioner
// (pymatting/preconditioner/ichol.py:6-181).
//
// Both halves of ichol are sequential by construction and run on the host,
// in one thread, in this library's own code: the factorisation (column j
// needs every earlier column that reaches row j) and the application
// z = L^-T L^-1 r, whose forward and back substitutions have dependency
// chains thousands of rows long (8734 levels for the 23632 unknowns of a
// 320x240 image). A level-scheduled GPU solve of such a factor measured about
// 11 us per level on MI355X, a dependent global-memory round trip per level,
// i.e. about 100 ms per solve, against 3.4 ms for the same solve on one host
// core. The application therefore copies r to pinned host memory, runs the
// two substitutions and copies z back, on the caller's stream; everything
// else of the PCG stays on the GPU.
//
// The substitutions sum a row's terms in ascending order with separate
// multiplies and subtractions (compiled with -ffp-contract=off): the forward
// one column by column and the back one row by row, as the reference's
// ichol.py:116-145 do, so the results are bit-identical to them.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "common.h"
#include "internal.h"

struct hm_ichol {
    long long n = 0, nnz = 0;
    // The factor in the reference's CSC layout (column j: the diagonal
    // first, then the rows below it in ascending order).
    std::vector<double> Lv;
    std::vector<int64_t> Lr, Lp;
    double* pinned = nullptr;  // host staging of r and z (n doubles)
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

// Forward substitution L y = x in place, column by column: y_j = x_j / L_jj,
// then x_i -= L_ij y_j for the rows below (ichol.py:116-129).
void forward(const hm_ichol* h, double* x) {
    const double* v = h->Lv.data();
    const int64_t* r = h->Lr.data();
    const int64_t* p = h->Lp.data();
    for (long long j = 0; j < h->n; ++j) {
        const double t = x[j] / v[p[j]];
        x[j] = t;
        for (int64_t k = p[j] + 1; k < p[j + 1]; ++k) x[r[k]] -= v[k] * t;
    }
}

// Back substitution L^T z = y in place, row by row from the last:
// z_i = (y_i - sum_k L_ki z_k) / L_ii over the rows k below i in column i
// (ichol.py:132-145).
void backward(const hm_ichol* h, double* x) {
    const double* v = h->Lv.data();
    const int64_t* r = h->Lr.data();
    const int64_t* p = h->Lp.data();
    for (long long i = h->n - 1; i >= 0; --i) {
        double s = x[i];
        for (int64_t k = p[i] + 1; k < p[i + 1]; ++k) s -= v[k] * x[r[k]];
        x[i] = s / v[p[i]];
    }
}

void release(hm_ichol* h) {
    if (!h) return;
    if (h->pinned) (void)hipHostFree(h->pinned);
    delete h;
}

}  // namespace

long long ichol_rows(const hm_ichol* h) { return h ? h->n : -1; }

// z = L^-T L^-1 r for device float64 vectors of h->n values: r is copied to
// pinned host memory on s, the substitutions run on the host, and z is copied
// back on s (the copy completes before the next application reuses the
// staging buffer, which it only touches after synchronising s). A finished
// solve (guard) is not applied again.
int ichol_apply(const hm_ichol* hc, const double* r, double* z, const CgState* guard,
                hipStream_t s) {
    hm_ichol* h = const_cast<hm_ichol*>(hc);
    const long long n = h->n;
    if (n == 0) return HM_OK;
    if (!h->pinned)
        HM_HIP(hipHostMalloc(reinterpret_cast<void**>(&h->pinned), sizeof(double) * n,
                             hipHostMallocDefault));
    HM_HIP(hipMemcpyAsync(h->pinned, r, sizeof(double) * n, hipMemcpyDeviceToHost, s));
    HM_HIP(hipStreamSynchronize(s));
    forward(h, h->pinned);
    backward(h, h->pinned);
    HM_HIP(hipMemcpyAsync(z, h->pinned, sizeof(double) * n, hipMemcpyHostToDevice, s));
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
    h->nnz = nnz;
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
    for (int64_t j = 0; j < n; ++j)
        for (int64_t k = Lp[j] + 1; k < Lp[j + 1]; ++k)
            if (Lr[k] <= j || Lr[k] >= n || (k > Lp[j] + 1 && Lr[k] <= Lr[k - 1])) {
                release(h);
                set_error("hm_ichol_create: column %lld is not strictly lower triangular "
                          "with ascending rows", (long long)j);
                return HM_ERR_VALUE;
            }
    *out = h;
    return HM_OK;
}

int64_t hm_ichol_nnz(const hm_ichol* h) { return h ? h->nnz : -1; }

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
