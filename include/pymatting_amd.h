/* pymatting_amd: C ABI of the MI355X (gfx950) closed-form matting library.
 *
 * Plain C types only: device pointers are `void*`/typed pointers obtained
 * from the caller's allocator (e.g. torch.Tensor.data_ptr()), streams are
 * `hipStream_t` passed as `void*`. Every function returns an int status
 * (HM_OK or a negative HM_ERR_*); hm_last_error() returns the message of the
 * last failure on the calling thread.
 *
 * The Python package pymatting_amd binds these with ctypes and mirrors the
 * reference's Python API (pymatting 1.1.15) on top of them.
 */
#ifndef PYMATTING_AMD_H
#define PYMATTING_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HM_OK 0
#define HM_ERR_ARG (-1)         /* bad argument (maps to ValueError / TypeError) */
#define HM_ERR_HIP (-2)         /* HIP runtime failure (RuntimeError) */
#define HM_ERR_NOCONV (-3)      /* CG did not converge within maxiter (ValueError) */
#define HM_ERR_UNSUPPORTED (-4) /* valid request this build does not implement */
#define HM_ERR_VALUE (-5)       /* invalid data, e.g. singular coarse matrix */
#define HM_ERR_CALLBACK (-6)    /* a host callback reported failure */

/* Precision of vectors and operator arithmetic. */
#define HM_F64 0
#define HM_F32 1

/* Operator forms of the level-0 closed-form operator. */
#define HM_OP_LAPLACIAN 0 /* L (windows with all pixels known skipped if a class map is given) */
#define HM_OP_LAMBDA 1    /* A = L + lambda * diag(known)  (make_linear_system) */
#define HM_OP_MASKED 2    /* [L_UU 0; 0 I]: hard constraints of estimate_alpha_cf */

/* Preconditioners. */
#define HM_PC_NONE 0
#define HM_PC_JACOBI 1
#define HM_PC_VCYCLE 2
#define HM_PC_CSR 3  /* z = M r with a matrix M (cg.py:59-62) */
#define HM_PC_HOST 4 /* z = M(r) through a host callback (callable M, cg.py:57-58) */
#define HM_PC_ICHOL 5 /* z = L^-T L^-1 r with an incomplete Cholesky factor (ichol.py:175-181);
                       * the substitutions run on the host, the state is read every iteration */

typedef struct hm_cf hm_cf;   /* device image + closed-form window operator */
typedef struct hm_dia hm_dia; /* stored 25-point grid stencil (coarse levels, ELL form) */
typedef struct hm_csr hm_csr; /* general CSR matrix (cg on arbitrary matrices) */
typedef struct hm_vc hm_vc;   /* V-cycle hierarchy */
typedef struct hm_ichol hm_ichol; /* incomplete Cholesky factor */
typedef struct hm_lkm hm_lkm;     /* large-kernel (box filter) matting Laplacian */

typedef struct hm_cg_stats {
    int iterations;     /* completed CG iterations */
    int converged;      /* 1 if the residual test passed */
    double norm_b;      /* ||b|| */
    double norm_r;      /* final recursive residual norm */
} hm_cg_stats;

const char* hm_last_error(void);
int hm_version(void);

/* Copies between host and device memory on a stream (kind: 0 = device to
 * host, 1 = host to device, 2 = device to device); synchronises for kind 0. */
int hm_copy(void* dst, const void* src, int64_t bytes, int kind, void* stream);

/* ---- trimap split: pymatting/util/util.py:439-525 (trimap_split) ----
 * cls[i] = 1 if trimap[i] >= fg_threshold, 2 if <= bg_threshold, else 0.
 * counts[0] = #fg, counts[1] = #bg (device int64[2]). */
int hm_trimap_classes(const double* trimap, int64_t n, double fg_threshold, double bg_threshold,
                      uint8_t* cls, int64_t* counts, void* stream);

/* ---- cf_laplacian: pymatting/laplacian/cf_laplacian.py:5-177 ----
 * image_hwc: device float64 (h, w, 3). cls: nullable device class map.
 * mode: HM_OP_*. radius must be 1. */
int hm_cf_create(const double* image_hwc, int h, int w, double epsilon, int radius,
                 const uint8_t* cls, double lambda_value, int mode, void* stream, hm_cf** out);
int hm_cf_destroy(hm_cf* op);
/* Reference CSR layout of cf_laplacian (values (n,5,5), indices, indptr),
 * entries bit-identical to the reference's serial assembly. */
int hm_cf_assemble_csr(hm_cf* op, double* values, int64_t* indices, int64_t* indptr,
                       void* stream);
/* y = A x with the matrix-free stencil; prec selects float64/float32 vectors. */
int hm_cf_apply(hm_cf* op, int prec, const void* x, void* y, void* stream);
/* Where the float32 matrix-free passes of this operator (hm_cf_apply with
 * HM_F32, float32 V-cycles and the mixed solves over it) take each 3x3
 * window's mean and inverse covariance from: 0 (default) the float64 values
 * computed once at creation (the quantities of cf_laplacian.py:38-92) rounded
 * to float32 and read per pass; 1 recomputed per pass from the image tile the
 * pass stages, in float64 arithmetic, then rounded to float32. float64 passes
 * always use the stored values. No counterpart in the reference, which only
 * assembles the matrix. */
int hm_cf_set_windows(hm_cf* op, int recompute);
/* d = diag(A) (float64). */
int hm_cf_diagonal(hm_cf* op, double* d, void* stream);
/* Right-hand side: lambda*is_fg (HM_OP_LAMBDA, make_linear_system,
 * laplacian.py:46-47) or -(L m)_U with m = is_fg on known pixels (HM_OP_MASKED,
 * estimate_alpha_cf.py:66-74). */
int hm_cf_rhs(hm_cf* op, double* b, void* stream);

/* Canonical CSR of the operator's matrix, as scipy leaves (L + diags(c)) after
 * tocsr() and sum_duplicates() (laplacian.py:49-55): in-image couplings with
 * a non-zero value, columns ascending. hm_cf_system_indptr fills indptr
 * (device int64[n+1]) and returns nnz through *nnz; hm_cf_system_fill then
 * writes indices and values. */
int hm_cf_system_indptr(hm_cf* op, int64_t* indptr, int64_t* nnz, void* stream);
int hm_cf_system_fill(hm_cf* op, const int64_t* indptr, int64_t* indices, double* values,
                      void* stream);

/* cf_laplacian.py:5-177 for any radius r >= 1: the reference's CSR with
 * (4r+1)^2 slots per row (values n*(4r+1)^2, indices likewise, indptr n+1,
 * device arrays), entries summed over windows in the reference's order.
 * known (nullable, device uint8, 1 = known) skips windows whose pixels are
 * all known (cf_laplacian.py:34-35). */
int hm_cf_laplacian_csr(const double* image_hwc, int h, int w, double epsilon, int radius,
                        const uint8_t* known, double* values, int64_t* indices, int64_t* indptr,
                        void* stream);

/* ---- lkm_laplacian: pymatting/laplacian/lkm_laplacian.py:6-76 ----
 * The closed-form Laplacian of (2r+1)^2 windows applied with box filters
 * (He et al. 2010), matrix-free. image_hwc: device float64 (h, w, 3), h and
 * w at least 2r + 1 (boxfilter.py:140-141). cls (nullable, device class
 * map, nonzero = known) makes the operator L + lambda diag(known), the
 * system of estimate_alpha_lkm.py:51-58. hm_lkm_apply: y = A x (the
 * reference's L_matvec, :51-65); hm_lkm_diagonal: diag(A) (:67-74), or its
 * inverse (the Jacobi preconditioner of :55-61) when inverse != 0;
 * hm_lkm_rhs: b = lambda * is_fg. They use the handle's work buffers:
 * calls on one handle must not overlap. */
int hm_lkm_create(const double* image_hwc, int h, int w, double epsilon, int radius,
                  const uint8_t* cls, double lambda_value, void* stream, hm_lkm** out);
int hm_lkm_apply(hm_lkm* L, const double* x, double* y, void* stream);
int hm_lkm_diagonal(hm_lkm* L, int inverse, double* d, void* stream);
int hm_lkm_rhs(hm_lkm* L, double* b, void* stream);
int hm_lkm_destroy(hm_lkm* L);

/* ---- stored 25-point operators (ELL/DIA, slot s = (dy+2)*5 + (dx+2)) ----
 * coef: device float64 [25][h*w] slot-major. The operator must be symmetric
 * (as every matrix handed to CG is): the library stores the 13 slots with
 * (dy, dx) >= (0, 0) and reads the others from the neighbouring rows.
 * hm_dia_coefficients writes the full [25][h*w] form back. */
int hm_dia_create(const double* coef, int h, int w, void* stream, hm_dia** out);
int hm_dia_from_cf(hm_cf* op, void* stream, hm_dia** out);
int hm_dia_destroy(hm_dia* A);
int hm_dia_apply(hm_dia* A, int prec, const void* x, void* y, void* stream);
int hm_dia_coefficients(hm_dia* A, double* coef, void* stream);

/* ---- general CSR (cg.py on arbitrary sparse/dense A) ---- */
int hm_csr_create(int64_t n, int64_t nnz, const int64_t* indptr, const int64_t* indices,
                  const double* values, hm_csr** out);
int hm_csr_destroy(hm_csr* A);
int hm_csr_apply(hm_csr* A, const double* x, double* y, void* stream);
/* The reduced system of estimate_alpha_cf.py:57-70 from a CSR Laplacian L
 * (device arrays, e.g. the reference layout of cf_laplacian) and the class
 * map cls (0 unknown, 1 foreground, 2 background): L_U = L[U, :][:, U] as
 * scipy's fancy indexing leaves it (the unknown rows in order; in each, the
 * entries whose column is unknown, renumbered by rank, in stored order,
 * duplicates and explicit zeros kept) and rhs = -R m with
 * R = L[U, :][:, K], m = is_fg[K]. hm_csr_reduced_indptr fills indptr_u
 * (device int64, room for n + 1) and returns n_u and nnz_u;
 * hm_csr_reduced_fill writes indices_u, values_u (nnz_u) and rhs_u (n_u).
 * hm_reduced_finalize writes alpha = clip(x, 0, 1) over all n pixels with
 * x = trimap on known pixels and x_u on unknown ones (:72-76). */
int hm_csr_reduced_indptr(int64_t n, const int64_t* indptr, const int64_t* indices,
                          const uint8_t* cls, int64_t* indptr_u, int64_t* n_u, int64_t* nnz_u,
                          void* stream);
int hm_csr_reduced_fill(int64_t n, const int64_t* indptr, const int64_t* indices,
                        const double* values, const uint8_t* cls, const int64_t* indptr_u,
                        int64_t* indices_u, double* values_u, double* rhs_u, void* stream);
int hm_reduced_finalize(int64_t n, const uint8_t* cls, const double* x_u, const double* trimap,
                        double* alpha, void* stream);
/* make_linear_system (laplacian.py:49-55) on an arbitrary CSR matrix A (device
 * arrays): the canonical CSR of A + diag(diag) that scipy leaves after
 * L + diags(c), tocsr() and sum_duplicates(): per row, equal columns summed
 * in stored order, diag[i] added on the diagonal, zero sums dropped, columns
 * ascending. hm_csr_add_diagonal_indptr fills out_indptr (n + 1) and
 * returns nnz; hm_csr_add_diagonal_fill writes the indices and values. Rows
 * may hold up to 2048 entries (HM_ERR_UNSUPPORTED beyond). */
int hm_csr_add_diagonal_indptr(int64_t n, const int64_t* indptr, const int64_t* indices,
                               const double* values, const double* diag, int64_t* out_indptr,
                               int64_t* nnz, void* stream);
int hm_csr_add_diagonal_fill(int64_t n, const int64_t* indptr, const int64_t* indices,
                             const double* values, const double* diag,
                             const int64_t* out_indptr, int64_t* out_indices, double* out_values,
                             void* stream);

/* ---- vcycle: pymatting/preconditioner/vcycle.py:6-153 ----
 * Geometric multigrid on the (h, w) grid of the operator: full-weighting
 * restriction, Galerkin coarse operators R A P, damped Jacobi smoothing and a
 * dense direct solve once h*w <= direct_solve_size. A grid above that size
 * whose coarse grid (h/2, w/2) would be empty (a 1-row grid, say) ends the
 * hierarchy: its cycle is num_pre_iter + num_post_iter damped Jacobi sweeps
 * from zero, which is what the reference's recursion into the empty grid
 * computes (that sum must then be positive). Exactly one of cf / dia is
 * non-null. prec: HM_F64, or HM_F32 for float32 coefficients and smoothing. */
int hm_vcycle_create(hm_cf* cf, hm_dia* dia, int h, int w, int num_pre_iter,
                     int num_post_iter, double omega, int direct_solve_size, int prec,
                     void* stream, hm_vc** out);
/* As hm_vcycle_create, with two extensions the reference does not have:
 * max_levels > 0 caps the number of grids (the coarsest is then the
 * max_levels-th, (h >> (max_levels-1), w >> (max_levels-1))); if that grid
 * has more than 2048 points it is solved approximately by coarse_sweeps
 * damped Jacobi sweeps from zero (coarse_sweeps = 0 then fails). With
 * max_levels = 0 and coarse_sweeps = 0 it is hm_vcycle_create. */
int hm_vcycle_create_ex(hm_cf* cf, hm_dia* dia, int h, int w, int num_pre_iter,
                        int num_post_iter, double omega, int direct_solve_size, int prec,
                        int max_levels, int coarse_sweeps, void* stream, hm_vc** out);
/* Direct-solve-only "V-cycle" for a dense n x n device matrix (n <= 2048):
 * the case h*w <= direct_solve_size of vcycle.py:59-60 for operators that are
 * not grid stencils. */
int hm_vcycle_create_dense(const double* A_dense, int n, int prec, void* stream, hm_vc** out);
int hm_vcycle_destroy(hm_vc* vc);
int hm_vcycle_levels(hm_vc* vc); /* number of grids including the direct level */
/* No counterpart in the reference: the launch structure of a cycle, a bit
 * mask. 1: consecutive coarse grids of the ascent (below 2^18 points, from
 * the third grid on, cycles with one pre- and one post-smoothing sweep) in
 * one launch; 2: pairs of those grids of the descent in one launch; 4: the
 * coarsest grids that one workgroup would run in LDS (the tail) applied as
 * their precomputed dense matrix. The default is 5. Bits 1 and 2 leave the
 * cycle bit-identical; bit 4 changes it by rounding only (the tail's sums in
 * another order). 0 runs every grid in a launch of its own. */
int hm_vcycle_set_launches(hm_vc* vc, int pairs);
/* z = V-cycle(r) (vectors in the hierarchy's precision). */
int hm_vcycle_apply(hm_vc* vc, const void* r, void* z, void* stream);

/* ---- cg: pymatting/solver/cg.py:4-117 ----
 * Host callbacks receive device pointers to float64 vectors of length n and
 * return 0 on success. hm_apply_cb computes y = f(x); hm_iter_cb is called
 * after every iteration's update with the current x and r, before the
 * convergence test, like cg.py's callback. */
typedef int (*hm_apply_cb)(void* ctx, const double* x, double* y);
typedef int (*hm_iter_cb)(void* ctx, const double* x, const double* r, double norm_b,
                          double norm_r);

typedef struct hm_cg_args {
    /* operator: exactly one of cf, dia, csr, lkm, a_cb (then n gives the size) */
    hm_cf* cf;
    hm_dia* dia;
    hm_csr* csr;
    hm_apply_cb a_cb;
    void* a_ctx;
    int64_t n;
    /* preconditioner */
    int pc_kind;              /* HM_PC_* */
    const double* inv_diag;   /* HM_PC_JACOBI: 1 / diag(A) */
    hm_vc* vc;                /* HM_PC_VCYCLE (float64 hierarchy) */
    hm_csr* pc_csr;           /* HM_PC_CSR */
    hm_apply_cb m_cb;         /* HM_PC_HOST */
    void* m_ctx;
    /* vectors (device float64) */
    const double* b;
    const double* x0;         /* nullable: start from zeros */
    double* x;
    double atol;
    double rtol;
    int maxiter;
    int reorthogonalize;
    int check_every;          /* iterations enqueued between host checks */
    hm_iter_cb callback;      /* nullable */
    void* cb_ctx;
    hm_ichol* ichol;          /* HM_PC_ICHOL */
    hm_lkm* lkm;              /* operator: large-kernel Laplacian */
} hm_cg_args;

/* Preconditioned CG. Returns HM_ERR_NOCONV after maxiter iterations without
 * convergence; stats are filled in either case. */
int hm_cg_solve(const hm_cg_args* args, hm_cg_stats* stats, void* stream);

/* Mixed precision: float32 stencil, SpMV and smoothing inside CG with float64
 * dot products, wrapped in float64 iterative refinement on the residual
 * b - A x (recomputed with the float64 operator). Exactly one of cf (matrix-
 * free) / dia (stored stencil) is the operator; vc32 must be HM_F32. */
int hm_cg_solve_mixed(hm_cf* cf, hm_dia* dia, hm_vc* vc32, const double* b, double* x,
                      double rtol, double inner_rtol, int max_outer, int maxiter,
                      hm_cg_stats* stats, void* stream);

/* Mixed precision without restarts: one CG whose recurrences (stencil, SpMV,
 * V-cycle) run in float32 with float64 dot products, the solution
 * accumulated in float64, and the recursive residual replaced by the true
 * float64 residual b - A x whenever it has dropped by a factor delta since
 * the last replacement; convergence (||b - A x|| < rtol ||b||) is confirmed
 * on a true residual. The state is read back every min(check_every, 8)
 * iterations (replacements are decided there). Same operator arguments as
 * hm_cg_solve_mixed. graph = 1 records an iteration without a replacement
 * (SpMV, update, V-cycle) once per alternation of the direction buffers into
 * a HIP graph and replays it, on a stream of the calling thread's own
 * ordered after and before `stream`; results are bit-identical to graph = 0
 * (direct launches). */
int hm_cg_solve_reliable(hm_cf* cf, hm_dia* dia, hm_vc* vc32, const double* b, double* x,
                         double rtol, double delta, int maxiter, int check_every, int graph,
                         hm_cg_stats* stats, void* stream);

/* ---- ichol: pymatting/preconditioner/ichol.py:6-301 ----
 * hm_ichol_factor: one attempt of the thresholded incomplete Cholesky
 * factorisation (_ichol, ichol.py:6-113) of a canonical CSC matrix given in
 * host memory (its lower triangle is used), with `shift` added to the
 * diagonal. *result receives the factor's nnz, or -1 when a pivot is not
 * positive, or -2 when more than max_nnz entries would be stored (the
 * reference's return codes; HM_OK is returned in all three cases and *out
 * is null for -1/-2). The factorisation runs on the host in one thread and
 * makes no device call. hm_ichol_create makes the handle of a given factor
 * (CholeskyDecomposition(Ltuple), ichol.py:148-162) from host CSC arrays
 * whose columns start with the diagonal, rows ascending. */
int hm_ichol_factor(int64_t n, const int64_t* indptr, const int64_t* indices,
                    const double* values, double discard_threshold, double shift,
                    int64_t max_nnz, double relative_discard_threshold, int diag_keep_discarded,
                    int64_t* result, hm_ichol** out);
int hm_ichol_create(int64_t n, const double* Lv, const int64_t* Lr, const int64_t* Lp,
                    hm_ichol** out);
int64_t hm_ichol_nnz(const hm_ichol* h);
/* Copies the factor (CSC: Lv[nnz], Lr[nnz], Lp[n+1]) to host arrays. */
int hm_ichol_copy_factor(const hm_ichol* h, double* Lv, int64_t* Lr, int64_t* Lp);
/* z = L^-T L^-1 r (CholeskyDecomposition.__call__, ichol.py:175-181) for
 * device float64 vectors: r is copied to pinned host memory on `stream`, the
 * forward and back substitutions run on the host (sequential chains of
 * thousands of rows; bit-identical to the reference's), and z is copied back
 * on `stream`. Calls on one handle must not overlap (it keeps the staging
 * buffer). */
int hm_ichol_apply(hm_ichol* h, const double* r, double* z, void* stream);
int hm_ichol_destroy(hm_ichol* h);

/* z = r * inv_diag (jacobi.py:24-31 applied to a vector). */
int hm_diag_apply(const double* inv_diag, const double* r, double* z, int64_t n, void* stream);

/* alpha = clip(x, 0, 1) with known pixels (cls != 0) replaced by their trimap
 * value first (estimate_alpha_cf.py:72-76). cls and trimap may be null. */
int hm_alpha_finalize(const double* x, const double* trimap, const uint8_t* cls, double* alpha,
                      int64_t n, void* stream);

/* ---- batched solve (no reference counterpart: pymatting solves one image
 * per call; this is cg.py:4-117 with the V-cycle of vcycle.py:46-100 run on
 * many independent systems at once) ----
 * ops[i]: the stored level-0 operator of image i; vcs[i]: its V-cycle, built
 * on ops[i] (float64 for precision HM_F64; float32 for the mixed form,
 * precision HM_F32), all with the same sweeps and omega and a direct
 * coarsest solve. level0_shape / coarse_shape: tile shape of the finest and
 * of the Galerkin grids (0: 64x8 points, two per thread; 1: 32x8, one per
 * thread). The batch keeps its own stream; hm_batch_solve orders it after
 * and before `stream`. Each image iterates until its own residual test
 * passes (per-image convergence masks); graph_iters > 0 records that many
 * iterations (rounded down to even) into a HIP graph that is replayed;
 * 0 launches them directly. rhs/x: host arrays of count device float64
 * vectors; stats: count entries. */
typedef struct hm_batch hm_batch;
int hm_batch_create(int count, hm_dia* const* ops, hm_vc* const* vcs, int precision,
                    int level0_shape, int coarse_shape, hm_batch** out);
int hm_batch_solve(hm_batch* b, const double* const* rhs, double* const* x, double rtol,
                   double inner_rtol, int max_outer, int maxiter, int graph_iters,
                   hm_cg_stats* stats, void* stream);
/* The mixed form with reliable updates (see hm_cg_solve_reliable) for a batch
 * created with precision HM_F32: one float32 CG per image with its solution in
 * float64; residual replacement per image after a drop by delta; every
 * image's convergence is confirmed on its float64 residual. Iterations run in
 * units of 8 between state read-backs. graph = 0 launches every kernel
 * directly; graph = 1 records each unit iteration's two fixed parts (SpMV,
 * finalize and update; the V-cycle) once into HIP graphs and replays them,
 * with the per-unit residual replacements launched directly between them.
 * Both give bit-identical results. */
int hm_batch_solve_reliable(hm_batch* b, const double* const* rhs, double* const* x,
                            double rtol, double delta, int maxiter, int graph,
                            hm_cg_stats* stats, void* stream);
int hm_batch_destroy(hm_batch* b);

/* ---- kernel boundary vs grid barrier (no reference counterpart) ----
 * Measures, on `stream`, for a grid of `workgroups` workgroups of 256 threads
 * (at most one per compute unit): the mean time per launch of `reps`
 * back-to-back empty launches (*launch_us), and per barrier of `reps`
 * grid-wide barriers inside one launch, cooperative groups' grid.sync()
 * (*coop_us) and an atomic counter barrier with a bounded spin (*counter_us;
 * *ok = 0 if a spin ran out). These are the two ways the solve could separate
 * its steps: launches per step, or one persistent kernel. */
int hm_time_sync(int workgroups, int reps, double* launch_us, double* coop_us,
                 double* counter_us, int* ok, void* stream);

/* ---- adjoint of the alpha solve (no reference counterpart: pymatting is
 * not differentiable; this is the image-gradient step of the adjoint method
 * for the closed-form system of cf_laplacian.py:5-129 / laplacian.py:46-55) ----
 * For x = A(I)^-1 b with A = L(I) + lambda diag(known) and lam = A^-1 g,
 * writes dl/dI = -lam^T (dA/dI) x as a device float64 (h, w, 3) array (added
 * to grad_hwc when accumulate != 0). lam and x are device float64 [n]. */
int hm_cf_adjoint_image(hm_cf* op, const double* lam, const double* x, double* grad_hwc,
                        int accumulate, void* stream);

/* ---- estimate_foreground_ml: pymatting/foreground/estimate_foreground_ml.py:61-253 ----
 * image (h0,w0,3) and alpha (h0,w0) device float32 (HM_F32) or float64
 * (HM_F64); F, B outputs of the same type. level_h/level_w: n_levels+1 host
 * ints. */
int hm_fb_ml(const void* image, const void* alpha, int h0, int w0, int n_levels,
             const int* level_h, const int* level_w, double regularization, int n_small,
             int n_big, int small_size, double gradient_weight, int prec, void* F, void* B,
             void* stream);

/* ---- estimate_foreground_cf: pymatting/foreground/estimate_foreground_cf.py:8-130 ----
 * The 2n x 2n system A = U^T U + [[S, 0], [0, S]] of the closed-form
 * foreground (U = [diag(a), diag(1 - a)], S = sum_k D_k^T diag(reg + |D_k a|)
 * D_k over the neighbour offsets (dx[k], dy[k]), k < n_dirs <= 8, with
 * clamped neighbours, :69-90) as the canonical CSR scipy leaves (columns
 * ascending, zeros dropped), entries bit-identical to the reference's.
 * alpha: device float64 (h, w). hm_fgcf_system_indptr fills indptr (2n + 1)
 * and returns nnz; hm_fgcf_system_fill writes the indices and values.
 * hm_fgcf_channel forms b = [a I_c; (1 - a) I_c] and x0 = [F0_c; B0_c] for
 * channel c of (h, w, channels) float64 images (:100-112); hm_fgcf_store
 * writes clip(x, 0, 1) back into channel c of F and B (:114-121). */
int hm_fgcf_system_indptr(const double* alpha, int h, int w, double regularization,
                          int n_dirs, const int* dx, const int* dy, int64_t* indptr,
                          int64_t* nnz, void* stream);
int hm_fgcf_system_fill(const double* alpha, int h, int w, double regularization, int n_dirs,
                        const int* dx, const int* dy, const int64_t* indptr, int64_t* indices,
                        double* values, void* stream);
int hm_fgcf_channel(const double* image, const double* alpha, const double* F0,
                    const double* B0, int64_t n, int channels, int channel, double* b, double* x0,
                    void* stream);
int hm_fgcf_store(const double* x, int64_t n, int channels, int channel, double* F, double* B,
                  void* stream);

/* ---- allocation-free foreground (no counterpart in the reference) ----
 * hm_fb_ml without any device allocation, for callers that cannot allocate
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc). Same arguments, kernels and results as hm_fb_ml; the
 * workspace is the library's static device array of 64 MiB per device
 * (kept until the process ends), which serves float32 images up to about
 * 0.8 Mpix (e.g. 1024x768) and float64 images up to about half that; larger
 * ones are rejected with HM_ERR_ARG before any device work. The workspace
 * is shared by every call on a device: calls that overlap on one device
 * race. */
int hm_fb_ml_noalloc(const void* image, const void* alpha, int h0, int w0, int n_levels,
                     const int* level_h, const int* level_w, double regularization,
                     int n_small, int n_big, int small_size, double gradient_weight, int prec,
                     void* F, void* B, void* stream);

/* ---- measurement (no reference counterpart) ----
 * hm_prof_set(1) starts recording HIP events around every launch of the
 * library's kernel classes (on the stream each kernel runs on) and discards
 * earlier records; hm_prof_query sums the recorded durations of one class
 * (HM_PROF_* below) for one precision (0 float64, 1 float32), with the number
 * of launches and the grid points they covered. */
#define HM_PROF_CF_APPLY 0
#define HM_PROF_CF_RESTRICT 1
#define HM_PROF_CF_SMOOTH 2
#define HM_PROF_CF_RESIDUAL 3
#define HM_PROF_DIA_APPLY 4     /* stored level-0 stencil */
#define HM_PROF_DIA_RESTRICT 5
#define HM_PROF_DIA_SMOOTH 6
#define HM_PROF_DIA_RESIDUAL 7
#define HM_PROF_CG_UPDATE 8
#define HM_PROF_CG_PUPDATE 9
#define HM_PROF_DENSE 10
#define HM_PROF_FG_SWEEP 11
#define HM_PROF_COARSE_APPLY 12    /* stored stencils of Galerkin levels >= 2 */
#define HM_PROF_COARSE_RESTRICT 13
#define HM_PROF_COARSE_SMOOTH 14
#define HM_PROF_COARSE_RESIDUAL 15
#define HM_PROF_CG_UPDATE_X64 16   /* float32 CG update with the solution in float64 */
#define HM_PROF_CG_UPDATE_R 17     /* CG update of r alone (x beside the V-cycle tail) */
#define HM_PROF_LEVEL1_APPLY 18    /* the first Galerkin level */
#define HM_PROF_LEVEL1_RESTRICT 19
#define HM_PROF_LEVEL1_SMOOTH 20
#define HM_PROF_LEVEL1_RESIDUAL 21
#define HM_PROF_CFREC_APPLY 22     /* matrix-free level 0, windows recomputed per pass */
#define HM_PROF_CFREC_RESTRICT 23
#define HM_PROF_CFREC_SMOOTH 24
#define HM_PROF_CFREC_RESIDUAL 25
int hm_prof_set(int enabled);
/* Mean duration (microseconds) of reps back-to-back launches of the finest
 * grid's pre-smooth + residual + restriction and of its prolongation +
 * post-smooth, on the hierarchy's own operator and buffers with r (device
 * vector in the hierarchy's precision) as the right-hand side: the two
 * fine-grid kernels of each V(1,1) cycle, timed with HIP events. */
int hm_vcycle_time_fine(hm_vc* vc, const void* r, int reps, double* restrict_us,
                        double* smooth_us, void* stream);
int hm_prof_query(int kind, int prec, double* total_ms, int64_t* launches, int64_t* points);

#ifdef __cplusplus
}
#endif

#endif /* PYMATTING_AMD_H */
