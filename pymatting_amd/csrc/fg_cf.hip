// Closed-form foreground and background estimation on gfx950
// (pymatting/foreground/estimate_foreground_cf.py:8-130).
//
// The system is A [F; B] = b per colour channel, 2n unknowns, with
//   A = U^T U + [[S, 0], [0, S]],  U = [diag(a), diag(1 - a)],
//   S = sum_k D_k^T diag(reg + |D_k a|) D_k,
// where D_k is the directional difference (D_k x)_i = x_i - x_{n_k(i)} to the
// clamped neighbour n_k(i) = clamp(i + (dx_k, dy_k)) (a zero row where the
// neighbour is the pixel itself), and b = [a I_c; (1 - a) I_c].
//
// k_fgcf_rows assembles A row by row in the canonical CSR scipy leaves
// (columns ascending, zero entries dropped) with the reference's arithmetic:
// each D_k^T W D_k entry sums its terms in ascending order of the row r of
// D_k they come from (+g_r on the diagonals of r and n_k(r), -g_r on the two
// couplings), the directions' matrices are added in list order, and U^T U's
// products are added before S on the diagonals. Every entry is therefore the
// same double as in the reference's scipy matrix (tested against the oracle).
#include "common.h"
#include "internal.h"

namespace hm {

namespace {

constexpr int kMaxDirs = 8;
constexpr int kMaxRowCols = 4 * kMaxDirs + 2;  // F columns of a row plus its B column

struct FgcfDirs {
    int k;
    int dx[kMaxDirs], dy[kMaxDirs];
};

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// Accumulator of one row's entries (column, value), folded with `+` in the
// order they arrive.
struct RowAcc {
    long long col[kMaxRowCols];
    double val[kMaxRowCols];
    int m = 0;
    bool overflow = false;
    __device__ __forceinline__ void add(long long c, double v) {
        for (int q = 0; q < m; ++q)
            if (col[q] == c) {
                val[q] = val[q] + v;
                return;
            }
        if (m == kMaxRowCols) {
            overflow = true;
            return;
        }
        col[m] = c;
        val[m] = v;
        ++m;
    }
};

// Row i of S: for each direction in order, that direction's entries of row i
// (diagonal: +g_r for the rows r in ascending order whose difference involves
// i; couplings: -g_i to n_k(i), -g_r to each r with n_k(r) = i), folded into
// the row.
__device__ void s_row(const double* __restrict__ a, int h, int w, double reg,
                      const FgcfDirs& dirs, long long i, RowAcc& S) {
    const int yi = (int)(i / w), xi = (int)(i - (long long)yi * w);
    for (int k = 0; k < dirs.k; ++k) {
        const int dx = dirs.dx[k], dy = dirs.dy[k];
        // Rows r whose clamped neighbour is i: per axis a contiguous range.
        auto range = [](int c, int d, int n_, int& lo, int& hi) {
            lo = n_;  // empty
            hi = -1;
            for (int r = 0; r < n_; ++r)  // n_ is an image side; r + d clamps to c
                if (clampi(r + d, 0, n_ - 1) == c) {
                    if (r < lo) lo = r;
                    hi = r;
                }
        };
        int x0, x1, y0, y1;
        // Bounded scans: only sides near the clamped borders can have more
        // than one row; interior pixels have the single r = i - d.
        if (xi > 0 && xi < w - 1) {
            x0 = x1 = xi - dx;
            if (x0 < 0 || x0 >= w || clampi(x0 + dx, 0, w - 1) != xi) { x0 = w; x1 = -1; }
        } else {
            range(xi, dx, w, x0, x1);
        }
        if (yi > 0 && yi < h - 1) {
            y0 = y1 = yi - dy;
            if (y0 < 0 || y0 >= h || clampi(y0 + dy, 0, h - 1) != yi) { y0 = h; y1 = -1; }
        } else {
            range(yi, dy, h, y0, y1);
        }
        const long long ni = (long long)clampi(yi + dy, 0, h - 1) * w + clampi(xi + dx, 0, w - 1);
        const bool own = ni != i;  // row i of D_k is not zero
        const double gi = own ? reg + fabs(a[i] - a[ni]) : 0.0;
        // Diagonal of S2_k: ascending r over {r: n_k(r) = i, r != i} and i.
        double diag = 0.0;
        bool have = false;
        bool own_done = !own;
        for (int yr = y0; yr <= y1; ++yr)
            for (int xr = x0; xr <= x1; ++xr) {
                const long long r = (long long)yr * w + xr;
                if (r == i) continue;
                if (!own_done && i < r) {
                    diag = have ? diag + gi : gi;
                    have = true;
                    own_done = true;
                }
                const long long nr = (long long)clampi(yr + dy, 0, h - 1) * w +
                                     clampi(xr + dx, 0, w - 1);
                const double gr = reg + fabs(a[r] - a[nr]);
                diag = have ? diag + gr : gr;
                have = true;
            }
        if (!own_done) {
            diag = have ? diag + gi : gi;
            have = true;
        }
        // Within one direction each coupling (i, j) has a single term, so
        // only the order of the directions matters for the fold.
        if (have) S.add(i, diag);
        if (own) S.add(ni, -gi);
        for (int yr = y0; yr <= y1; ++yr)
            for (int xr = x0; xr <= x1; ++xr) {
                const long long r = (long long)yr * w + xr;
                if (r == i) continue;
                const long long nr = (long long)clampi(yr + dy, 0, h - 1) * w +
                                     clampi(xr + dx, 0, w - 1);
                S.add(r, -(reg + fabs(a[r] - a[nr])));
            }
    }
}

// mode 0: counts[row] for the 2n rows; mode 1: the entries from ptr.
__global__ void k_fgcf_rows(const double* __restrict__ a, int h, int w, double reg,
                            FgcfDirs dirs, int mode, int64_t* __restrict__ counts,
                            const int64_t* __restrict__ ptr, int64_t* __restrict__ idx,
                            double* __restrict__ val, int* __restrict__ overflow) {
    const long long n = (long long)h * w;
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 2 * n) return;
    const bool brow = t >= n;
    const long long i = brow ? t - n : t;
    RowAcc S;
    s_row(a, h, w, reg, dirs, i, S);
    if (S.overflow) {
        atomicOr(overflow, 1);
        return;
    }
    const double ai = a[i], bi = 1.0 - a[i];
    // Row F_i: U^T U gives a_i^2 at (F_i, F_i) and a_i (1 - a_i) at (F_i, B_i);
    // row B_i: (1 - a_i) a_i at (B_i, F_i) and (1 - a_i)^2 at (B_i, B_i).
    const double uu_diag = brow ? bi * bi : ai * ai;
    const double uu_off = brow ? bi * ai : ai * bi;
    const long long off_col = brow ? i : n + i;
    const long long shift = brow ? n : 0;
    long long col[kMaxRowCols + 1];
    double v[kMaxRowCols + 1];
    int m = 0;
    for (int q = 0; q < S.m; ++q) {
        double x = S.val[q];
        if (S.col[q] == i && uu_diag != 0.0) x = uu_diag + x;
        col[m] = S.col[q] + shift;
        v[m] = x;
        ++m;
    }
    bool diag_seen = false;
    for (int q = 0; q < S.m; ++q) diag_seen |= S.col[q] == i;
    if (!diag_seen && uu_diag != 0.0) {
        col[m] = i + shift;
        v[m] = uu_diag;
        ++m;
    }
    if (uu_off != 0.0) {
        col[m] = off_col;
        v[m] = uu_off;
        ++m;
    }
    // Ascending columns (insertion sort of a short row), zeros dropped.
    for (int p = 1; p < m; ++p)
        for (int q = p; q > 0 && col[q - 1] > col[q]; --q) {
            const long long tc = col[q];
            col[q] = col[q - 1];
            col[q - 1] = tc;
            const double tv = v[q];
            v[q] = v[q - 1];
            v[q - 1] = tv;
        }
    if (mode == 0) {
        int cnt = 0;
        for (int q = 0; q < m; ++q) cnt += v[q] != 0.0 ? 1 : 0;
        counts[t] = cnt;
        return;
    }
    int64_t p = ptr[t];
    for (int q = 0; q < m; ++q)
        if (v[q] != 0.0) {
            idx[p] = col[q];
            val[p] = v[q];
            ++p;
        }
}

// b = [a I_c; (1 - a) I_c] and x0 = [F0_c; B0_c] for channel c of (h, w, d)
// images.
__global__ void k_fgcf_channel(const double* __restrict__ image, const double* __restrict__ a,
                               const double* __restrict__ F0, const double* __restrict__ B0,
                               long long n, int d, int c, double* __restrict__ b,
                               double* __restrict__ x0) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double I = image[d * i + c];
    b[i] = a[i] * I;
    b[n + i] = (1.0 - a[i]) * I;
    x0[i] = F0[d * i + c];
    x0[n + i] = B0[d * i + c];
}

// F[:, :, c] = clip(x[:n], 0, 1), B[:, :, c] = clip(x[n:], 0, 1).
__global__ void k_fgcf_store(const double* __restrict__ x, long long n, int d, int c,
                             double* __restrict__ F, double* __restrict__ B) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double f = x[i], g = x[n + i];
    F[d * i + c] = f < 0.0 ? 0.0 : (f > 1.0 ? 1.0 : f);
    B[d * i + c] = g < 0.0 ? 0.0 : (g > 1.0 ? 1.0 : g);
}

int make_dirs(int k, const int* dx, const int* dy, FgcfDirs& d) {
    if (k < 1 || k > kMaxDirs || !dx || !dy) {
        set_error("estimate_foreground_cf: 1 to %d neighbour offsets are supported, got %d",
                  kMaxDirs, k);
        return HM_ERR_ARG;
    }
    d.k = k;
    for (int q = 0; q < k; ++q) {
        d.dx[q] = dx[q];
        d.dy[q] = dy[q];
    }
    return HM_OK;
}

}  // namespace

}  // namespace hm

extern "C" {

int hm_fgcf_system_indptr(const double* alpha, int h, int w, double regularization,
                          int n_dirs, const int* dx, const int* dy, int64_t* indptr,
                          int64_t* nnz, void* stream) {
    using namespace hm;
    if (!alpha || !indptr || !nnz || h < 1 || w < 1) {
        set_error("hm_fgcf_system_indptr: invalid argument");
        return HM_ERR_ARG;
    }
    FgcfDirs d;
    HM_TRY(make_dirs(n_dirs, dx, dy, d));
    hipStream_t s = static_cast<hipStream_t>(stream);
    const long long n2 = 2LL * h * w;
    void* counts = nullptr;
    void* flag = nullptr;
    HM_TRY(dev_alloc(&counts, sizeof(int64_t) * n2));
    if (dev_alloc(&flag, sizeof(int)) != HM_OK) {
        dev_free(counts);
        return HM_ERR_HIP;
    }
    int over = 0, rc = HM_OK;
    hipError_t e = hipMemsetAsync(flag, 0, sizeof(int), s);
    if (e == hipSuccess) {
        k_fgcf_rows<<<div_up(n2, 256), 256, 0, s>>>(alpha, h, w, regularization, d, 0,
                                                    static_cast<int64_t*>(counts), nullptr,
                                                    nullptr, nullptr, static_cast<int*>(flag));
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(&over, flag, sizeof(int), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) {
        set_error("hm_fgcf_system_indptr: %s", hipGetErrorString(e));
        rc = HM_ERR_HIP;
    } else if (over) {
        set_error("estimate_foreground_cf: a row couples more pixels than supported");
        rc = HM_ERR_UNSUPPORTED;
    } else {
        rc = exclusive_offsets(static_cast<int64_t*>(counts), indptr, n2, nnz, s);
    }
    dev_free(counts);
    dev_free(flag);
    return rc;
}

int hm_fgcf_system_fill(const double* alpha, int h, int w, double regularization, int n_dirs,
                        const int* dx, const int* dy, const int64_t* indptr, int64_t* indices,
                        double* values, void* stream) {
    using namespace hm;
    if (!alpha || !indptr || !indices || !values || h < 1 || w < 1) {
        set_error("hm_fgcf_system_fill: invalid argument");
        return HM_ERR_ARG;
    }
    FgcfDirs d;
    HM_TRY(make_dirs(n_dirs, dx, dy, d));
    void* flag = nullptr;
    HM_TRY(dev_alloc(&flag, sizeof(int)));
    hipStream_t s = static_cast<hipStream_t>(stream);
    const long long n2 = 2LL * h * w;
    hipError_t e = hipMemsetAsync(flag, 0, sizeof(int), s);
    if (e == hipSuccess) {
        k_fgcf_rows<<<div_up(n2, 256), 256, 0, s>>>(alpha, h, w, regularization, d, 1, nullptr,
                                                    indptr, indices, values,
                                                    static_cast<int*>(flag));
        e = hipGetLastError();
    }
    dev_free(flag);
    if (e != hipSuccess) {
        set_error("hm_fgcf_system_fill: %s", hipGetErrorString(e));
        return HM_ERR_HIP;
    }
    return HM_OK;
}

int hm_fgcf_channel(const double* image, const double* alpha, const double* F0,
                    const double* B0, int64_t n, int channels, int channel, double* b, double* x0,
                    void* stream) {
    using namespace hm;
    if (!image || !alpha || !F0 || !B0 || !b || !x0 || n < 1 || channel < 0 ||
        channel >= channels) {
        set_error("hm_fgcf_channel: invalid argument");
        return HM_ERR_ARG;
    }
    k_fgcf_channel<<<div_up(n, 256), 256, 0, static_cast<hipStream_t>(stream)>>>(
        image, alpha, F0, B0, n, channels, channel, b, x0);
    HM_LAUNCH_CHECK();
    return HM_OK;
}

int hm_fgcf_store(const double* x, int64_t n, int channels, int channel, double* F, double* B,
                  void* stream) {
    using namespace hm;
    if (!x || !F || !B || n < 1 || channel < 0 || channel >= channels) {
        set_error("hm_fgcf_store: invalid argument");
        return HM_ERR_ARG;
    }
    k_fgcf_store<<<div_up(n, 256), 256, 0, static_cast<hipStream_t>(stream)>>>(x, n, channels,
                                                                              channel, F, B);
    HM_LAUNCH_CHECK();
    return HM_OK;
}

}  // extern "C"
