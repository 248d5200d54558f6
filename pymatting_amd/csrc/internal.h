// Host-side object layouts and launcher declarations shared by the library's
// translation units.
#pragma once

#include <vector>

#include "common.h"

// Device image + closed-form window operator (level 0 of every solve).
struct hm_cf {
    int h = 0, w = 0;
    long long n = 0;
    double eps = 1e-7;
    double lambda = 0.0;
    int mode = HM_OP_LAPLACIAN;
    int skip_known = 0;
    double* I64 = nullptr;   // planar [3][n]
    float* I32 = nullptr;    // planar [3][n]
    uint8_t* cls = nullptr;  // owned copy of the class map, or null
    double* wc = nullptr;    // exact per-window data [10][n]: mean, inverse covariance, flag
    float* wc32 = nullptr;   // float32 copy of the first 9 planes
    double* diag64 = nullptr;
    float* diag32 = nullptr;
    double* idiag64 = nullptr;  // 1 / diag (float32 copy rounded from float64)
    float* idiag32 = nullptr;
    int ready = 0;  // kCf* parts computed so far (cf_prepare)
    int recompute = 0;  // float32 passes recompute the windows from the image
};

// Stored symmetric 25-point stencil (slot-major [13][n], see engine.h) with
// its diagonal.
struct hm_dia {
    int h = 0, w = 0;
    long long n = 0;
    double* c64 = nullptr;
    float* c32 = nullptr;
    double* diag64 = nullptr;
    float* diag32 = nullptr;
    double* idiag64 = nullptr;  // 1 / diag (float32 copy rounded from float64)
    float* idiag32 = nullptr;
};

struct hm_ichol;  // ichol.hip
struct hm_lkm;    // lkm.hip

struct hm_csr {
    long long n = 0, nnz = 0;
    int64_t* indptr = nullptr;
    int64_t* indices = nullptr;
    double* values = nullptr;
};

namespace hm {

// One grid of the V-cycle hierarchy. Level 0 borrows the caller's operator;
// coarser levels own a Galerkin stencil. Vectors are in the hierarchy's
// precision (double or float, stored as raw bytes).
struct VcLevel {
    int h = 0, w = 0;
    long long n = 0;
    hm_dia* A = nullptr;   // owned for levels >= 1 (or level 0 given as DIA: borrowed)
    bool owns_A = false;
    void* b = nullptr;     // right-hand side (restricted residual); level 0 uses the input
    void* x = nullptr;     // correction / smoothed iterate
    void* t = nullptr;     // scratch for extra Jacobi sweeps
    void* u = nullptr;     // scratch for extra Jacobi sweeps
};

}  // namespace hm

struct hm_vc {
    int prec = HM_F64;
    int pre = 1, post = 1;
    double omega = 0.8;
    int direct_size = 64;
    hm_cf* cf = nullptr;  // level-0 operator when matrix-free
    std::vector<hm::VcLevel> levels;  // last level is solved directly
    double* dense_inv64 = nullptr;    // inverse of the coarsest matrix (n_c x n_c)
    float* dense_inv32 = nullptr;
    int n_direct = 0;
    // > 0: the coarsest grid is solved approximately by this many damped
    // Jacobi sweeps from zero (a fixed polynomial in D^-1 A, so the cycle
    // stays a symmetric preconditioner) instead of the dense inverse.
    int coarse_sweeps = 0;
    hm::Reducer red{};
    // The tail's constant LDS contents (stencils of its LDS-resident grids,
    // then the dense inverse when staged) packed in their LDS order, in this
    // hierarchy's precision; built once for the tail starting at tail_start.
    void* tail_pack = nullptr;
    size_t tail_start = 0;
    // The tail as a dense row-major tail_rows x tail_cols matrix, applied
    // instead of the tail's launch when present: from the right-hand side of
    // grid tail_src (the tail's top grid, or the grid above it with its
    // restriction folded in) to the top grid's correction.
    void* tail_mat = nullptr;
    int tail_rows = 0, tail_cols = 0;
    size_t tail_src = 0;
    // Launch structure of the coarse levels, a bit mask: 1 runs consecutive
    // levels of the ascent in one launch, 2 pairs of levels of the descent
    // (pair.h), 4 applies the tail as its dense matrix (vcycle.h); 0 runs
    // every level in launches of its own. 1 and 2 leave the cycle
    // bit-identical, 4 changes its rounding.
    int pairs = 5;
};

namespace hm {

// cf_exact.hip
int cf_upload_image(const double* hwc, hm_cf* op, hipStream_t s);
int cf_exact_setup(hm_cf* op, hipStream_t s);
// Parts of a matrix-free operator computed on first need: the diagonal and
// its inverse (Jacobi smoothing, jacobi()), and the float32 copies of the
// image and window data (float32 stencil passes).
constexpr int kCfDiag = 1, kCfF32 = 2;
int cf_prepare(hm_cf* op, int need, hipStream_t s);
int cf_assemble(hm_cf* op, double* values, long long stride_i, long long stride_s, int s_lo,
                int64_t* indices, int64_t* indptr, int mode, hipStream_t s);
// The stored symmetric stencil of op's mode-adjusted operator with its
// float32 copy, diagonal and inverse diagonal (A allocated at op's shape).
int cf_assemble_sym(hm_cf* op, hm_dia* A, hipStream_t s);
int cf_rhs(hm_cf* op, double* b, hipStream_t s);
int cf_system_counts(hm_cf* op, int64_t* counts, hipStream_t s);
int cf_system_fill(hm_cf* op, const int64_t* indptr, int64_t* indices, double* values,
                   hipStream_t s);

// scan.hip: out[0] = 0, out[k + 1] = sum(in[0..k]); returns the total.
int exclusive_offsets(const int64_t* in, int64_t* out, long long n, int64_t* total,
                      hipStream_t s);
int trimap_classes(const double* trimap, long long n, double fg, double bg, uint8_t* cls,
                   int64_t* counts, hipStream_t s);

// grid.hip
int dia_galerkin(const hm_dia* fine, hm_dia* coarse, hipStream_t s);  // + dia_finish's outputs
int dia_finish(hm_dia* A, hipStream_t s);  // diagonal + float32 copies
int dia_pack(const double* full, hm_dia* A, hipStream_t s);     // [25][n] -> [13][n]
int dia_unpack(const hm_dia* A, double* full, hipStream_t s);   // [13][n] -> [25][n]
constexpr int kMaxDirect = 2048;  // largest coarsest grid solved by the dense inverse
// status (a device int, nullable): left set when the matrix is singular, for
// the caller to read at its next synchronisation; null reads it back here.
int dense_inverse(const hm_dia* A, double* inv64, float* inv32, hipStream_t s,
                  int* status = nullptr);
int dense_inverse_from_dense(const double* A, int n, double* inv64, float* inv32, hipStream_t s);
int csr_apply(const hm_csr* A, const double* x, double* y, CgState* st, Reducer* red, int kind,
              hipStream_t s);

// ichol.hip: z = L^-T L^-1 r with the factor's triangular solves.
int ichol_apply(const hm_ichol* h, const double* r, double* z, const CgState* guard,
                hipStream_t s);
long long ichol_rows(const hm_ichol* h);
// lkm.hip: y = A x with the large-kernel operator; its size.
int lkm_apply(const hm_lkm* L, const double* x, double* y, hipStream_t s);
long long lkm_rows(const hm_lkm* L);

// Stage two of a reduction over `nblocks` partials (engine.h k_finalize).
int launch_finalize(int nvals, const Reducer& red, unsigned nblocks, CgState* st, int kind,
                    const CgState* guard, hipStream_t s);

// Raises a kernel's dynamic LDS limit on the current device (once per
// kernel and device, thread-safe).
int set_max_dynamic_lds(const void* kernel, int bytes);

// Device memory helpers (hipMalloc-backed).
int dev_alloc(void** p, size_t bytes);
void dev_free(void* p);
void pool_trim();  // returns cached blocks to HIP
int make_reducer(Reducer* r, int max_blocks, int nvals);
void free_reducer(Reducer* r);

}  // namespace hm
