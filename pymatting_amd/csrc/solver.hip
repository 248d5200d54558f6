// Closed-form operators, V-cycle and PCG drivers, and the C ABI.
//
// The PCG follows pymatting/solver/cg.py:4-117 with all scalars on the
// device: the SpMV kernel's last block forms alpha = rz / <p, Ap>, the update
// kernel's last block tests ||r|| < atol or ||r|| < rtol ||b|| and raises a
// `done` flag that turns every later kernel of the solve into a no-op, and
// the preconditioner's last kernel forms beta = (1 / rz_old) * rz_new. The
// host enqueues `check_every` iterations at a time and reads the state once
// per chunk.
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <utility>

#include "common.h"
#include "engine.h"
#include "internal.h"
#include "vcycle.h"
#include "pair.h"

namespace {
thread_local std::string g_last_error;
}

namespace hm {

void set_error(const char* fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_last_error = buf;
}

// Mixed-precision helpers.
__global__ void k_scale_to_f32(const double* __restrict__ a, double scale, float* __restrict__ b,
                               long long n) {
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = (float)(a[i] * scale);
}

__global__ void k_axpy_f32_to_f64(double* __restrict__ x, double scale,
                                  const float* __restrict__ e, long long n) {
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] += scale * (double)e[i];
}

// z = V-cycle(r) (vcycle.py:46-100). If st is given, the final kernel also
// forms <r, z> and finalises it as `kind`, through `red` (the caller's
// reducer, which a deferred CG update shares) or the V-cycle's own. `side`
// (only when vc_has_tail) is work for the tail's launch to carry.
template <typename T>
int vcycle_apply_t(hm_vc* vc, const T* r, T* z, CgState* st, const Reducer* red, int kind,
                   const CgState* guard, hipStream_t s, const TailSide* side) {
    const size_t L = vc->levels.size();
    if (side && side->x && !vc_has_tail(vc)) {
        set_error("internal: tail side work for a V-cycle without a tail");
        return HM_ERR_ARG;
    }
    const T omega = (T)vc->omega;
    const Reducer& R = red ? *red : vc->red;
    const DotSink nodot{nullptr, Reducer{}, 0};
    const DotSink dot{st, R, kind};
    if (L == 1 && vc->coarse_sweeps > 0) {
        // The finest grid cannot be coarsened (see vc_create): its pre- and
        // post-smoothing, coarse_sweeps damped Jacobi sweeps from zero on r,
        // ping-ponging through .x and .t so that the last one lands in z and
        // forms <r, z>.
        const LevelOp<T> op = vc_level_op<T>(vc, 0);
        T* bufs[2] = {vptr<T>(vc->levels[0].x), vptr<T>(vc->levels[0].t)};
        const int S = vc->coarse_sweeps;
        const T* prev = nullptr;
        for (int k = 1; k <= S; ++k) {
            T* outp = (k == S) ? z : bufs[(S - 1 - k) % 2];
            const OutSmooth<T> sm{r, op.iD, omega, outp, k == S ? dot : nodot};
            if (k == 1)
                HM_TRY(run_level<T>(op, InZero<T>{}, sm, guard, s));
            else
                HM_TRY(run_level<T>(op, InVec<T>{prev}, sm, guard, s));
            prev = outp;
        }
        return HM_OK;
    }
    if (L == 1) {
        k_dense_apply<T><<<1, kBlock, 0, s>>>(vc_inv<T>(vc), r, z, vc->n_direct, st, R, guard);
        HM_LAUNCH_CHECK();
        if (st) return launch_finalize(1, R, 1, st, kind, guard, s);
        return HM_OK;
    }
    std::vector<T*> result(L, nullptr), pre_x(L, nullptr);
    const size_t tail = vc_tail_start(vc);
    // With the tail applied as its matrix, the matrix may start at the grid
    // above the tail (its restriction folded in): the descent stops there.
    size_t tail_src = tail;
    const bool tail_mat = tail < L - 1 && vc->coarse_sweeps == 0 &&
                          vc_tail_matrix(vc, tail, &tail_src);
    const size_t desc_end = tail_mat ? tail_src : tail;
    // Coarse levels that run several to a launch (pair.h).
    const VcPlan plan = vc_plan(vc, tail, desc_end);
    // Descend: pre-smooth, residual, restrict.
    for (size_t l = 0; l + 1 < L && l < desc_end; ++l) {
        if (plan.down_pair[l]) {
            HM_TRY(launch_pair_down<T>(vc, l, guard, s));
            ++l;  // level l + 1 restricted too
            continue;
        }
        const VcLevel& lev = vc->levels[l];
        const VcLevel& cl = vc->levels[l + 1];
        const LevelOp<T> op = vc_level_op<T>(vc, l);
        const T* b = (l == 0) ? r : vptr<T>(lev.b);
        const OutRestrict<T> rest{b, vptr<T>(cl.b), cl.w, cl.h, nullptr, nodot};
        if (vc->pre == 0) {
            HM_TRY(run_level<T>(op, InZero<T>{}, rest, guard, s));
        } else if (vc->pre == 1) {
            HM_TRY(run_level<T>(op, InJac0<T>{b, op.iD, omega}, rest, guard, s));
        } else {
            T* cur = vptr<T>(lev.x);
            T* nxt = vptr<T>(lev.t);
            HM_TRY(run_level<T>(op, InJac0<T>{b, op.iD, omega},
                                OutSmooth<T>{b, op.iD, omega, cur, nodot}, guard, s));
            for (int k = 2; k < vc->pre; ++k) {
                HM_TRY(run_level<T>(op, InVec<T>{cur}, OutSmooth<T>{b, op.iD, omega, nxt, nodot},
                                    guard, s));
                std::swap(cur, nxt);
            }
            HM_TRY(run_level<T>(op, InVec<T>{cur}, rest, guard, s));
            pre_x[l] = cur;
        }
    }
    // Coarse tail in one launch (its top grid's correction lands in .x), or
    // the direct solve alone (vcycle.py:59-60).
    if (vc->coarse_sweeps > 0) {
        // Approximate coarsest solve: x = omega b / D, then damped Jacobi
        // sweeps, ping-ponging so that the last one lands in .x.
        const VcLevel& cl = vc->levels[L - 1];
        const LevelOp<T> op = vc_level_op<T>(vc, L - 1);
        const T* cb = vptr<T>(cl.b);
        T* xs = vptr<T>(cl.x);
        T* ts = vptr<T>(cl.t);
        const int m = vc->coarse_sweeps - 1;  // sweeps after the first iterate
        if (m == 0) {
            k_jac0<T><<<div_up(cl.n, 256), 256, 0, s>>>(cb, op.iD, omega, xs, cl.n, guard);
            HM_LAUNCH_CHECK();
        } else {
            T* cur = (m % 2 == 1) ? xs : ts;
            T* oth = (m % 2 == 1) ? ts : xs;
            HM_TRY(run_level<T>(op, InJac0<T>{cb, op.iD, omega},
                                OutSmooth<T>{cb, op.iD, omega, cur, nodot}, guard, s));
            for (int k = 1; k < m; ++k) {
                HM_TRY(run_level<T>(op, InVec<T>{cur}, OutSmooth<T>{cb, op.iD, omega, oth, nodot},
                                    guard, s));
                std::swap(cur, oth);
            }
        }
    } else if (tail < L - 1) {
        HM_TRY(launch_tail<T>(vc, tail, guard, s, side));
    } else {
        const VcLevel& cl = vc->levels[L - 1];
        k_dense_apply<T><<<1, kBlock, 0, s>>>(vc_inv<T>(vc), vptr<T>(cl.b), vptr<T>(cl.x),
                                              vc->n_direct, nullptr, Reducer{}, guard);
        HM_LAUNCH_CHECK();
    }
    result[tail] = vptr<T>(vc->levels[tail].x);
    // Ascend: prolongate the correction and post-smooth (vcycle.py:95-100).
    for (size_t l = tail; l-- > 0;) {
        if (plan.up_start[l] >= 0) {  // levels l, l - 1, ..., a in one launch
            const size_t a = (size_t)plan.up_start[l];
            T* za = vptr<T>(vc->levels[a].x);
            HM_TRY(launch_up_chain<T>(vc, a, l - a + 1, result[l + 1], za, guard, s));
            result[a] = za;
            l = a;
            continue;
        }
        const VcLevel& lev = vc->levels[l];
        const VcLevel& cl = vc->levels[l + 1];
        const LevelOp<T> op = vc_level_op<T>(vc, l);
        const T* b = (l == 0) ? r : vptr<T>(lev.b);
        const bool top = (l == 0);
        T* avail[2];
        int na = 0;
        for (void* cand : {lev.x, lev.t, lev.u})
            if (cand != nullptr && vptr<T>(cand) != pre_x[l] && na < 2) avail[na++] = vptr<T>(cand);
        if (!top && na < 1) {
            set_error("internal: missing V-cycle work vectors");
            return HM_ERR_ARG;
        }
        if (vc->post >= 2 && na < (top ? 1 : 2)) {
            set_error("internal: missing V-cycle work vectors");
            return HM_ERR_ARG;
        }
        T* target = top ? z : avail[0];
        T* other = top ? (na > 0 ? avail[0] : nullptr) : (na > 1 ? avail[1] : nullptr);
        const DotSink fin = (top && st) ? dot : nodot;
        const T* e = result[l + 1];
        int rc;
        if (vc->pre == 0)
            rc = post_stage<T>(op, InZero<T>{}, b, e, cl.h, cl.w, vc->post, omega, target, other,
                               fin, r, guard, s, R);
        else if (vc->pre == 1)
            rc = post_stage<T>(op, InJac0<T>{b, op.iD, omega}, b, e, cl.h, cl.w, vc->post, omega,
                               target, other, fin, r, guard, s, R);
        else
            rc = post_stage<T>(op, InVec<T>{pre_x[l]}, b, e, cl.h, cl.w, vc->post, omega,
                               target, other, fin, r, guard, s, R);
        HM_TRY(rc);
        result[l] = target;
    }
    return HM_OK;
}

int vcycle_apply(hm_vc* vc, const void* r, void* z, CgState* st, const Reducer* red, int kind,
                 const CgState* guard, hipStream_t s, const TailSide* side = nullptr) {
    if (vc->prec == HM_F64)
        return vcycle_apply_t<double>(vc, static_cast<const double*>(r), static_cast<double*>(z),
                                      st, red, kind, guard, s, side);
    return vcycle_apply_t<float>(vc, static_cast<const float*>(r), static_cast<float*>(z), st,
                                 red, kind, guard, s, side);
}

}  // namespace hm

// ===================================================================== PCG

namespace hm {

struct OpHandle {
    const hm_cf* cf = nullptr;
    const hm_dia* dia = nullptr;
    const hm_csr* csr = nullptr;
    const hm_lkm* lkm = nullptr;  // large-kernel operator (float64 only)
    hm_apply_cb cb = nullptr;  // host callback operator (float64 only)
    void* cb_ctx = nullptr;
    long long cb_n = 0;
    long long n() const {
        return cf ? cf->n : (dia ? dia->n : (csr ? csr->n : (lkm ? lkm_rows(lkm) : cb_n)));
    }
    int h() const { return cf ? cf->h : (dia ? dia->h : 1); }
    int w() const { return cf ? cf->w : (dia ? dia->w : (int)n()); }
};

struct PcHandle {
    int kind = HM_PC_NONE;
    const double* inv_diag = nullptr;
    hm_vc* vc = nullptr;
    const hm_csr* csr = nullptr;
    const hm_ichol* ichol = nullptr;
    hm_apply_cb cb = nullptr;
    void* cb_ctx = nullptr;
};

int reducer_blocks(long long n, int h, int w) {
    long long a = (long long)div_up(w, 64) * div_up(h, 8);
    long long b = vec_blocks(n);
    long long c = div_up(n, kBlock);
    long long m = a > b ? a : b;
    m = m > c ? m : c;
    return (int)(m + 64);
}

// Runs a host callback on device vectors: the stream is drained first so the
// callback sees finished data, and the callback's own copies complete before
// it returns.
inline int host_call(hm_apply_cb cb, void* ctx, const double* x, double* y, hipStream_t s) {
    HM_HIP(hipStreamSynchronize(s));
    if (cb(ctx, x, y) != 0) {
        set_error("host callback failed");
        return HM_ERR_CALLBACK;
    }
    return HM_OK;
}

// float32 passes of a matrix-free operator read its float32 copies.
template <typename T>
inline int cf_ready(const OpHandle& A, hipStream_t s) {
    if (A.cf && sizeof(T) == 4) return cf_prepare(const_cast<hm_cf*>(A.cf), kCfF32, s);
    return HM_OK;
}

template <typename T>
int op_apply(const OpHandle& A, const T* x, T* y, const DotSink& ds, const CgState* guard,
             hipStream_t s) {
    HM_TRY(cf_ready<T>(A, s));
    if (A.cf) return run_level<T>(cf_level<T>(A.cf), InVec<T>{x}, OutApply<T>{y, ds}, guard, s);
    if (A.dia)
        return run_level<T>(dia_level<T>(A.dia), InVec<T>{x}, OutApply<T>{y, ds}, guard, s);
    if constexpr (std::is_same<T, double>::value) {
        if (A.csr) {
            Reducer red = ds.red;
            return csr_apply(A.csr, x, y, ds.st, ds.st ? &red : nullptr, ds.kind, s);
        }
        if (A.lkm)
            HM_TRY(lkm_apply(A.lkm, x, y, s));
        else
            HM_TRY(host_call(A.cb, A.cb_ctx, x, y, s));
        if (ds.st) {
            const long long n = A.n();
            HM_TRY(launch_dot<double>(x, y, nullptr, n, ds.st, ds.red, ds.kind, s));
        }
        return HM_OK;
    }
    set_error("CSR and callback operators are float64 only");
    return HM_ERR_UNSUPPORTED;
}

// q = A p with p = z + beta p_old formed on the fly (stencil operators), or a
// separate direction update followed by the product (CSR / callbacks).
template <typename T>
int op_apply_pupdate(const OpHandle& A, const T* z, const T* p_old, T* p_new, T* q,
                     const DotSink& ds, CgState* st, hipStream_t s) {
    const InPupdate<T> in{z, p_old, st};
    OutApply<T> out{q, ds, p_new};
    HM_TRY(cf_ready<T>(A, s));
    if (A.cf) return run_level<T>(cf_level<T>(A.cf), in, out, st, s);
    if (A.dia) return run_level<T>(dia_level<T>(A.dia), in, out, st, s);
    const long long n = A.n();
    // p_new = p_old * beta + z as a pointwise pass, then the product.
    k_copy<T><<<div_up(n, 256), 256, 0, s>>>(p_old, p_new, n, st);
    HM_LAUNCH_CHECK();
    k_cg_pupdate<T><<<vec_blocks(n), kBlock, 0, s>>>(p_new, z, n, st);
    HM_LAUNCH_CHECK();
    return op_apply<T>(A, p_new, q, ds, st, s);
}

__global__ void k_sub_from(const double* __restrict__ b, double* __restrict__ r, long long n) {
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) r[i] = b[i] - r[i];
}

// r = b - A x.
template <typename T>
int op_residual(const OpHandle& A, const T* b, const T* x, T* r, hipStream_t s) {
    const DotSink nodot{nullptr, Reducer{}, 0};
    HM_TRY(cf_ready<T>(A, s));
    if (A.cf)
        return run_level<T>(cf_level<T>(A.cf), InVec<T>{x}, OutResidual<T>{b, r, nodot}, nullptr,
                            s);
    if (A.dia)
        return run_level<T>(dia_level<T>(A.dia), InVec<T>{x}, OutResidual<T>{b, r, nodot},
                            nullptr, s);
    if constexpr (std::is_same<T, double>::value) {
        HM_TRY(op_apply<double>(A, x, r, nodot, nullptr, s));
        k_sub_from<<<div_up(A.n(), 256), 256, 0, s>>>(b, r, A.n());
        HM_LAUNCH_CHECK();
        return HM_OK;
    }
    set_error("CSR and callback operators are float64 only");
    return HM_ERR_UNSUPPORTED;
}

template <typename T>
int pc_apply(const PcHandle& M, const T* r, T* z, CgState* st, const Reducer& red, int kind,
             const CgState* guard, long long n, hipStream_t s, const TailSide* side = nullptr) {
    if (M.kind == HM_PC_VCYCLE) return vcycle_apply(M.vc, r, z, st, &red, kind, guard, s, side);
    if (side && side->x) {
        set_error("internal: tail side work without a V-cycle");
        return HM_ERR_ARG;
    }
    if (M.kind == HM_PC_CSR || M.kind == HM_PC_HOST || M.kind == HM_PC_ICHOL) {
        if constexpr (std::is_same<T, double>::value) {
            if (M.kind == HM_PC_CSR)
                HM_TRY(csr_apply(M.csr, r, z, nullptr, nullptr, 0, s));
            else if (M.kind == HM_PC_ICHOL)
                HM_TRY(ichol_apply(M.ichol, r, z, guard, s));
            else
                HM_TRY(host_call(M.cb, M.cb_ctx, r, z, s));
            if (st) {
                HM_TRY(launch_dot<double>(r, z, nullptr, n, st, red, kind, s));
            }
            return HM_OK;
        }
        set_error("matrix, ichol and callback preconditioners are float64 only");
        return HM_ERR_UNSUPPORTED;
    }
    k_diag_pc<T><<<vec_blocks(n), kBlock, 0, s>>>(r, M.kind == HM_PC_JACOBI ? M.inv_diag : nullptr,
                                                  z, n, st, red, kind);
    HM_LAUNCH_CHECK();
    if (st) HM_TRY(launch_finalize(1, red, vec_blocks(n), st, kind, guard, s));
    return HM_OK;
}

// Device copy of the solve state plus a pinned host mirror (cached per
// thread: hipHostMalloc is slow).
struct DeviceState {
    CgState* dev = nullptr;
    CgState* host = nullptr;
    int init() {
        static thread_local CgState* pinned = nullptr;
        if (!pinned) HM_HIP(hipHostMalloc(reinterpret_cast<void**>(&pinned), sizeof(CgState) * 16,
                                          hipHostMallocDefault));
        static thread_local int depth = 0;
        if (depth >= 16) {
            set_error("solver nesting too deep");
            return HM_ERR_ARG;
        }
        host = pinned + depth++;
        depth_ = &depth;
        HM_TRY(dev_alloc(reinterpret_cast<void**>(&dev), sizeof(CgState)));
        std::memset(host, 0, sizeof(CgState));
        return HM_OK;
    }
    ~DeviceState() {
        dev_free(dev);
        if (depth_) --*depth_;
    }
    int push(hipStream_t s) {
        HM_HIP(hipMemcpyAsync(dev, host, sizeof(CgState), hipMemcpyHostToDevice, s));
        return HM_OK;
    }
    int pull(hipStream_t s) {
        HM_HIP(hipMemcpyAsync(host, dev, sizeof(CgState), hipMemcpyDeviceToHost, s));
        HM_HIP(hipStreamSynchronize(s));
        return HM_OK;
    }
    int* depth_ = nullptr;
};

struct DevBuf {
    void* p = nullptr;
    ~DevBuf() { dev_free(p); }
    int alloc(size_t bytes) { return dev_alloc(&p, bytes); }
    template <typename T>
    T* as() const { return static_cast<T*>(p); }
};

struct PcgOptions {
    double atol = 0.0, rtol = 1e-7;
    int maxiter = 10000;
    int reorth = 0;
    int check_every = 8;
    hm_iter_cb callback = nullptr;
    void* cb_ctx = nullptr;
};

// The CG update with the alpha finalize folded in (cg.py:90-100): each of
// kUpdBlocks workgroups sums the SpMV's <p, A p> partials in the fixed order
// of k_finalize (so all agree with it and with each other), forms
// alpha = rz / pq, and updates x and r over its grid-stride share with
// ||r||^2 partials; workgroup 0 records pq and alpha in the state. x is
// float64 in the reliable-update solve (TX = double, T = float), else T.
// With kUpdX false only r is updated: x += alpha p then runs in the
// V-cycle tail's launch (TailSide), with the alpha recorded here.
constexpr int kUpdBlocks = 256;

template <typename TX, typename T, bool kUpdX = true>
__global__ __launch_bounds__(kFinBlock) void k_cg_update_pq(TX* __restrict__ x,
                                                            const T* __restrict__ p,
                                                            T* __restrict__ r,
                                                            const T* __restrict__ q, long long n,
                                                            CgState* st, Reducer red_pq,
                                                            unsigned pq_blocks, Reducer red) {
    if (solve_done(st)) return;
    constexpr int U = 4;
    const long long stride = (long long)gridDim.x * kFinBlock;
    const long long first = (long long)blockIdx.x * kFinBlock + threadIdx.x;
    TX xv[U];
    T pv[U], rv[U], qv[U];
    const auto load = [&](long long i0) {
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const long long i = i0 + u * stride;
            const bool in = i < n;
            if constexpr (kUpdX) {
                xv[u] = in ? x[i] : TX(0);
                pv[u] = in ? p[i] : T(0);
            }
            rv[u] = in ? r[i] : T(0);
            qv[u] = in ? q[i] : T(0);
        }
    };
    // The first chunk's loads go out before the partial sums of alpha, so
    // both latencies overlap.
    load(first);
    __shared__ double s_alpha;
    {
        const double* src[1] = {red_pq.partials};
        const unsigned cnt[1] = {pq_blocks};
        double sum[1];
        finalize_sums<1>(src, cnt, sum);
        if (threadIdx.x == 0) {
            if (blockIdx.x == 0) dot_finalize(st, kDotPQ, sum[0]);
            s_alpha = st->rz / sum[0];
        }
        __syncthreads();
    }
    const double alpha = s_alpha;
    const T alphaT = (T)alpha;
    double acc = 0.0;
    for (long long i0 = first; i0 < n; i0 += U * stride) {
        if (i0 != first) load(i0);
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const long long i = i0 + u * stride;
            if (i >= n) continue;
            if constexpr (!kUpdX) {
            } else if constexpr (std::is_same<TX, T>::value) {
                x[i] = xv[u] + alphaT * pv[u];
            } else {
                x[i] = xv[u] + alpha * (TX)pv[u];
            }
            const T ri = rv[u] - alphaT * qv[u];
            r[i] = ri;
            acc += (double)ri * (double)ri;
        }
    }
    double v[1] = {acc};
    block_partials<kFinBlock, 1>(v, red);
}

// Length of the next batch of iterations enqueued between two read-backs of
// the state: up to where the measured rate (||r||^2 ratio rho2 per
// iteration) predicts ||r||^2 < conv, between 1 and 2K iterations, or K
// while no rate is known. A batch that ends at the converging iteration
// leaves no frozen iterations behind it, and long batches keep read-backs
// from interrupting the stream of launches.
inline int next_batch(int K, double rr, double rho2, double conv) {
    if (!(rho2 > 0.0 && rho2 < 1.0)) return K;
    if (!(rr > conv)) return 1;
    const double j = std::ceil(std::log(conv / rr) / std::log(rho2));
    return (int)std::max(1.0, std::min(j, 2.0 * K));
}

// Preconditioned CG (cg.py:4-117) on device vectors of type T with float64
// scalars. Returns HM_ERR_NOCONV if maxiter iterations did not converge;
// stats are filled in either case. With host callbacks (callable A or M, or a
// per-iteration callback) the state is read after every iteration.
template <typename T>
int pcg_t(const OpHandle& A, const PcHandle& M, const T* b, const T* x0, T* x,
          const PcgOptions& opt, hm_cg_stats* stats, hipStream_t s) {
    const long long n = A.n();
    const int maxiter = opt.maxiter < 0 ? 0 : opt.maxiter;
    DevBuf r, p, p2, q, z, rold;
    HM_TRY(r.alloc(sizeof(T) * n));
    HM_TRY(p.alloc(sizeof(T) * n));
    HM_TRY(p2.alloc(sizeof(T) * n));
    HM_TRY(q.alloc(sizeof(T) * n));
    HM_TRY(z.alloc(sizeof(T) * n));
    if (opt.reorth) HM_TRY(rold.alloc(sizeof(T) * n));
    Reducer red{};
    const unsigned vb = vec_blocks(n);
    HM_TRY(make_reducer(&red, std::max<int>({reducer_blocks(n, A.h(), A.w()), (int)vb, kUpdBlocks}),
                        3));
    struct RedGuard { Reducer* r; ~RedGuard() { free_reducer(r); } } rg{&red};
    red.upd_blocks = vb;
    Reducer red_upd = red;  // slot 2: ||r||^2 partials of the CG update
    red_upd.partials += 2 * (size_t)red.max_blocks;
    DeviceState st;
    HM_TRY(st.init());
    st.host->atol = opt.atol;
    st.host->rtol = opt.rtol;
    HM_TRY(st.push(s));
    // ichol's application runs on the host (it synchronises the stream
    // anyway): the state is read after every iteration so that no
    // application follows convergence.
    const bool interactive =
        A.cb || M.kind == HM_PC_HOST || M.kind == HM_PC_ICHOL || opt.callback;

    // norm_b = ||b|| (cg.py:66). With host callbacks it is read here; the
    // other solves finalise it and the initial residual test on the device
    // (kDotNormBOnly, kDotInitRes: a passing test freezes the solve) and read
    // both with the first batch's state, with no round trip before it.
    double norm_b = 0.0, norm_r = 0.0;
    if (interactive) {
        HM_TRY(launch_dot<T>(b, b, nullptr, n, st.dev, red, kDotAux, s));
        HM_TRY(st.pull(s));
        norm_b = std::sqrt(st.host->aux);
    } else {
        HM_TRY(launch_dot<T>(b, b, nullptr, n, st.dev, red, kDotNormBOnly, s));
    }
    // x = x0 or 0; r = b - A x (cg.py:64-71)
    if (x0) {
        k_copy<T><<<div_up(n, 256), 256, 0, s>>>(x0, x, n, nullptr);
        HM_LAUNCH_CHECK();
        HM_TRY(op_residual<T>(A, b, x, r.as<T>(), s));
    } else {
        k_fill<T><<<div_up(n, 256), 256, 0, s>>>(x, T(0), n);
        HM_LAUNCH_CHECK();
        if (A.cb) {  // the reference evaluates A(x) even for x = 0
            HM_TRY(op_residual<T>(A, b, x, r.as<T>(), s));
        } else {
            k_copy<T><<<div_up(n, 256), 256, 0, s>>>(b, r.as<T>(), n, nullptr);
            HM_LAUNCH_CHECK();
        }
    }
    stats->iterations = 0;
    stats->converged = 0;
    if (interactive) {
        HM_TRY(launch_dot<T>(r.as<T>(), r.as<T>(), nullptr, n, st.dev, red, kDotAux, s));
        HM_TRY(st.pull(s));
        norm_r = std::sqrt(st.host->aux);
        stats->norm_b = norm_b;
        stats->norm_r = norm_r;
        if (norm_r < opt.atol || norm_r < opt.rtol * norm_b) {  // cg.py:75-76
            stats->converged = 1;
            return HM_OK;
        }
        st.host->norm_b = norm_b;
        st.host->rr = norm_r * norm_r;
        HM_TRY(st.push(s));
    } else {
        HM_TRY(launch_dot<T>(r.as<T>(), r.as<T>(), nullptr, n, st.dev, red, kDotInitRes, s));
        if (maxiter == 0) {  // no batch: read the initial test here
            HM_TRY(st.pull(s));
            stats->norm_b = st.host->norm_b;
            stats->norm_r = std::sqrt(st.host->rr);
            if (st.host->done) {
                stats->converged = 1;
                return HM_OK;
            }
            set_error("Conjugate gradient descent did not converge within %d iterations",
                      maxiter);
            return HM_ERR_NOCONV;
        }
    }
    // z = M r; rz = <r, z> (cg.py:78-80). p = z comes from the first fused
    // direction update with beta = 0 and a zero previous direction.
    HM_TRY(pc_apply<T>(M, r.as<T>(), z.as<T>(), st.dev, red, kDotRzFirst, nullptr, n, s));
    k_fill<T><<<div_up(n, 256), 256, 0, s>>>(p.as<T>(), T(0), n);
    HM_LAUNCH_CHECK();
    T* p_cur = p.as<T>();   // previous direction
    T* p_next = p2.as<T>(); // direction being formed

    const int check_every = interactive ? 1 : (opt.check_every < 1 ? 1 : opt.check_every);
    // Without a host round trip per iteration, the convergence test of the
    // update is finalised together with the next <r, z> (one launch fewer per
    // iteration); the preconditioner then runs once more after convergence,
    // which changes neither x nor the iteration count.
    const bool defer = !interactive && !opt.reorth;
    // Stencil operators in that mode also leave the alpha finalize to the
    // update kernel, which sums the SpMV's partials itself.
    const bool fused = defer && (A.cf || A.dia);
    if (fused) red.upd_blocks = kUpdBlocks;
    // In that mode, with an LDS tail in the V-cycle, x += alpha p runs in the
    // tail's launch on the CUs the tail leaves idle; the update kernel only
    // updates r.
    const bool x_in_tail = fused && M.kind == HM_PC_VCYCLE && vc_has_tail(M.vc);
    red_upd.upd_blocks = red.upd_blocks;
    int enqueued = 0;
    DotSink pq{st.dev, red, kDotPQ};
    pq.defer = fused;
    const unsigned pq_blocks =
        fused ? tile_map(OutApply<T>{}, A.h(), A.w(), A.h() / 2, A.w() / 2).ntiles : 0;
    int kb = check_every;  // length of the next batch (next_batch once a rate is known)
    double rr_prev = norm_r * norm_r, rho2 = 0.0;  // (non-interactive: from the first read)
    int it_prev = 0;
    while (enqueued < maxiter) {
        const int k = std::min(maxiter - enqueued, kb);
        for (int it = 0; it < k; ++it) {
            if (opt.reorth) {
                k_copy<T><<<div_up(n, 256), 256, 0, s>>>(r.as<T>(), rold.as<T>(), n, st.dev);
                HM_LAUNCH_CHECK();
            }
            HM_TRY(op_apply_pupdate<T>(A, z.as<T>(), p_cur, p_next, q.as<T>(), pq, st.dev, s));
            std::swap(p_cur, p_next);
            {
                ProfScope prof_u(x_in_tail ? kProfCgUpdateR : kProfCgUpdate, n, sizeof(T) == 4,
                                 s);
                if (x_in_tail)
                    k_cg_update_pq<T, T, false><<<kUpdBlocks, kFinBlock, 0, s>>>(
                        nullptr, p_cur, r.as<T>(), q.as<T>(), n, st.dev, red, pq_blocks, red_upd);
                else if (fused)
                    k_cg_update_pq<T, T><<<kUpdBlocks, kFinBlock, 0, s>>>(
                        x, p_cur, r.as<T>(), q.as<T>(), n, st.dev, red, pq_blocks, red_upd);
                else
                    k_cg_update<T><<<vb, kBlock, 0, s>>>(x, p_cur, r.as<T>(), q.as<T>(), n,
                                                         st.dev, red_upd);
                HM_LAUNCH_CHECK();
                if (!defer) HM_TRY(launch_finalize(1, red_upd, vb, st.dev, kDotUpdate, st.dev, s));
            }
            if (interactive) {
                HM_TRY(st.pull(s));
                if constexpr (std::is_same<T, double>::value) {
                    if (opt.callback &&
                        opt.callback(opt.cb_ctx, x, r.as<T>(), norm_b, std::sqrt(st.host->rr)) != 0) {
                        set_error("iteration callback failed");
                        return HM_ERR_CALLBACK;
                    }
                }
                if (st.host->done || st.host->nonfinite) break;
            }
            if (opt.reorth) {
                HM_TRY(pc_apply<T>(M, r.as<T>(), z.as<T>(), nullptr, red, 0, st.dev, n, s));
                HM_TRY(launch_dot<T>(r.as<T>(), z.as<T>(), rold.as<T>(), n, st.dev, red, kDotRzUpdate, s));
            } else {
                const TailSide side{x, p_cur, n, st.dev, std::is_same<T, double>::value ? 1 : 0};
                HM_TRY(pc_apply<T>(M, r.as<T>(), z.as<T>(), st.dev, red,
                                   defer ? kDotUpdateRz : kDotRzUpdate, st.dev, n, s,
                                   x_in_tail ? &side : nullptr));
            }
        }
        enqueued += k;
        HM_TRY(st.pull(s));
        if (!interactive && enqueued == k) {  // the first read: ||b||, initial ||r||^2
            norm_b = st.host->norm_b;
            rr_prev = st.host->rr0;
        }
        if (st.host->done || st.host->nonfinite) break;
        if (!interactive) {
            const double rr = st.host->rr;
            const int its = st.host->iters;
            if (its > it_prev && rr < rr_prev)
                rho2 = std::pow(rr / rr_prev, 1.0 / (its - it_prev));
            rr_prev = rr;
            it_prev = its;
            const double tol = std::max(opt.atol, opt.rtol * norm_b);
            kb = next_batch(check_every, rr, rho2, tol * tol);
        }
    }
    stats->iterations = st.host->iters;
    stats->norm_r = std::sqrt(st.host->rr);
    stats->norm_b = st.host->norm_b;
    stats->converged = st.host->done;
    if (!st.host->done) {
        set_error("Conjugate gradient descent did not converge within %d iterations", maxiter);
        return HM_ERR_NOCONV;
    }
    return HM_OK;
}

// Mixed precision: float32 inner PCG (float32 stencil, SpMV and V-cycle
// smoothing; float64 dot products) inside float64 iterative refinement:
//   r = b - A x (float64), e = A32^-1 (r / ||r||), x += ||r|| e.
int pcg_mixed(hm_cf* cf, hm_dia* dia, hm_vc* vc32, const double* b, double* x, double rtol,
              double inner_rtol, int max_outer, int maxiter, hm_cg_stats* stats, hipStream_t s) {
    const long long n = cf ? cf->n : dia->n;
    const int oh = cf ? cf->h : dia->h, ow = cf ? cf->w : dia->w;
    DevBuf r64, r32, e32;
    HM_TRY(r64.alloc(sizeof(double) * n));
    HM_TRY(r32.alloc(sizeof(float) * n));
    HM_TRY(e32.alloc(sizeof(float) * n));
    Reducer red{};
    HM_TRY(make_reducer(&red, reducer_blocks(n, oh, ow), 2));
    struct RedGuard { Reducer* r; ~RedGuard() { free_reducer(r); } } rg{&red};
    DeviceState st;
    HM_TRY(st.init());
    HM_TRY(st.push(s));
    const unsigned vb = vec_blocks(n);
    OpHandle A;
    A.cf = cf;
    A.dia = dia;
    PcHandle M;
    M.kind = HM_PC_VCYCLE;
    M.vc = vc32;

    HM_TRY(launch_dot<double>(b, b, nullptr, n, st.dev, red, kDotAux, s));
    HM_TRY(st.pull(s));
    const double norm_b = std::sqrt(st.host->aux);
    k_fill<double><<<div_up(n, 256), 256, 0, s>>>(x, 0.0, n);
    HM_LAUNCH_CHECK();
    k_copy<double><<<div_up(n, 256), 256, 0, s>>>(b, r64.as<double>(), n, nullptr);
    HM_LAUNCH_CHECK();
    int total = 0;
    stats->norm_b = norm_b;
    stats->converged = 0;
    for (int outer = 0; outer <= max_outer; ++outer) {
        HM_TRY(launch_dot<double>(r64.as<double>(), r64.as<double>(), nullptr, n, st.dev, red, kDotAux, s));
        HM_TRY(st.pull(s));
        const double norm_r = std::sqrt(st.host->aux);
        stats->norm_r = norm_r;
        if (norm_r < rtol * norm_b) {
            stats->converged = 1;
            break;
        }
        if (outer == max_outer || total >= maxiter || !std::isfinite(norm_r)) break;
        k_scale_to_f32<<<div_up(n, 256), 256, 0, s>>>(r64.as<double>(), 1.0 / norm_r,
                                                      r32.as<float>(), n);
        HM_LAUNCH_CHECK();
        hm_cg_stats inner{};
        PcgOptions io;
        io.atol = 0.0;
        io.rtol = inner_rtol;
        io.maxiter = maxiter - total;
        int rc = pcg_t<float>(A, M, r32.as<float>(), nullptr, e32.as<float>(), io, &inner, s);
        if (rc != HM_OK && rc != HM_ERR_NOCONV) return rc;
        total += inner.iterations;
        k_axpy_f32_to_f64<<<div_up(n, 256), 256, 0, s>>>(x, norm_r, e32.as<float>(), n);
        HM_LAUNCH_CHECK();
        HM_TRY(op_residual<double>(A, b, x, r64.as<double>(), s));
    }
    stats->iterations = total;
    if (!stats->converged) {
        set_error("Conjugate gradient descent did not converge within %d iterations", maxiter);
        return HM_ERR_NOCONV;
    }
    return HM_OK;
}

// Mixed precision with reliable updates (one CG, no restarts): the CG
// recurrences run in float32 (stencil, SpMV, V-cycle; float64 dot products)
// while the solution accumulates in float64 (x += alpha p), and every time
// the recursive residual has dropped by `delta` since the last replacement it
// is replaced by the true float64 residual b - A x (rounded to float32)
// before the preconditioner runs; convergence is declared on a true residual.

// Records ||r||^2 of a replacement iteration (the true residual's).
__global__ void k_note_rr(CgState* st) {
    if (!solve_done(st)) st->aux = st->rr;
}

// r32 = (float)r64 with ||r64||^2 partials (replacing the update's; the same
// kUpdBlocks-workgroup grid, so the partial counts match).
__global__ __launch_bounds__(kFinBlock) void k_replace_r(const double* __restrict__ r64,
                                                         float* __restrict__ r32, long long n,
                                                         const CgState* st, Reducer red) {
    if (solve_done(st)) return;
    double acc = 0.0;
    const long long stride = (long long)gridDim.x * kFinBlock;
    for (long long i = (long long)blockIdx.x * kFinBlock + threadIdx.x; i < n; i += stride) {
        const double v = r64[i];
        r32[i] = (float)v;
        acc += v * v;
    }
    double v[1] = {acc};
    block_partials<kFinBlock, 1>(v, red);
}

int pcg_reliable(hm_cf* cf, hm_dia* dia, hm_vc* vc32, const double* b, double* x, double rtol,
                 double delta, int maxiter, int check_every, bool graph, hm_cg_stats* stats,
                 hipStream_t s) {
    OpHandle A;
    A.cf = cf;
    A.dia = dia;
    PcHandle M;
    M.kind = HM_PC_VCYCLE;
    M.vc = vc32;
    const long long n = A.n();
    DevBuf r64, r, p, p2, q, z;
    HM_TRY(r64.alloc(sizeof(double) * n));
    HM_TRY(r.alloc(sizeof(float) * n));
    HM_TRY(p.alloc(sizeof(float) * n));
    HM_TRY(p2.alloc(sizeof(float) * n));
    HM_TRY(q.alloc(sizeof(float) * n));
    HM_TRY(z.alloc(sizeof(float) * n));
    Reducer red{};
    HM_TRY(make_reducer(&red, std::max<int>(reducer_blocks(n, A.h(), A.w()), kUpdBlocks), 3));
    struct RedGuard { Reducer* r; ~RedGuard() { free_reducer(r); } } rg{&red};
    red.upd_blocks = kUpdBlocks;
    Reducer red_upd = red;
    red_upd.partials += 2 * (size_t)red.max_blocks;
    DeviceState st;
    HM_TRY(st.init());
    HM_TRY(st.push(s));
    stats->iterations = 0;
    stats->converged = 0;
    // x = 0, r = (float)b. ||b|| is finalised on the device (kDotNormB, which
    // also freezes the solve for b = 0, solved by x = 0) and reaches the host
    // with the first batch's state: no round trip before the first batch.
    std::memset(st.host, 0, sizeof(CgState));
    st.host->atol = 0.0;
    st.host->rtol = rtol;
    HM_TRY(st.push(s));
    HM_TRY(launch_dot<double>(b, b, nullptr, n, st.dev, red, kDotNormB, s));
    double norm_b = 0.0;  // set from the first read-back
    k_fill<double><<<div_up(n, 256), 256, 0, s>>>(x, 0.0, n);
    HM_LAUNCH_CHECK();
    k_replace_r<<<kUpdBlocks, kFinBlock, 0, s>>>(b, r.as<float>(), n, nullptr, red_upd);
    HM_LAUNCH_CHECK();
    k_fill<float><<<div_up(n, 256), 256, 0, s>>>(p.as<float>(), 0.0f, n);
    HM_LAUNCH_CHECK();
    HM_TRY(pc_apply<float>(M, r.as<float>(), z.as<float>(), st.dev, red, kDotRzFirst, nullptr, n,
                           s));
    float* p_cur = p.as<float>();
    float* p_next = p2.as<float>();
    // With an LDS tail in the V-cycle, x += alpha p runs in the tail's launch
    // on the CUs the tail leaves idle, and the update kernel only updates r;
    // replacement iterations need x first and update both.
#ifndef HM_TAIL_SIDE
#define HM_TAIL_SIDE 1
#endif
    const bool x_beside_tail = HM_TAIL_SIDE && vc_has_tail(vc32);
    // The state is read back every K iterations. A replacement is placed at
    // the iteration where the residual norm is predicted (from the rate of
    // the last batch) to have dropped by delta since the last replacement;
    // deciding this only at read-backs far apart lets the float32 recursion
    // drift from the true residual and stall the convergence.
#ifndef HM_RELIABLE_BATCH
#define HM_RELIABLE_BATCH 8
#endif
    constexpr int kReliableBatch = HM_RELIABLE_BATCH;
    const int K = check_every < 1 ? kReliableBatch : std::min(check_every, kReliableBatch);
    // <p, A p> partials of the SpMV, summed by the update kernel.
    DotSink pq{st.dev, red, kDotPQ};
    pq.defer = true;
    const unsigned pq_blocks = tile_map(OutApply<float>{}, A.h(), A.w(), A.h() / 2, A.w() / 2).ntiles;
    double rr_replaced = 0.0;  // ||r||^2 at the last replacement (from the first read-back)
    double rr_prev = 0.0, rho2 = 0.0;  // rate: ||r||^2 ratio per iteration
    int it_prev = 0;   // iterations at the last read-back
    int rep_idx = -1;  // replacement index in the next batch (none before a rate is known)
    int enqueued = 0;
    // The launches of an iteration without a replacement: the fused SpMV,
    // the update and the V-cycle (pc_cur -> pc_next).
    auto plain_iteration = [&](float* pc_cur, float* pc_next) -> int {
        HM_TRY(op_apply_pupdate<float>(A, z.as<float>(), pc_cur, pc_next, q.as<float>(), pq,
                                       st.dev, s));
        if (x_beside_tail) {
            ProfScope prof_u(kProfCgUpdateR, n, 1, s);
            k_cg_update_pq<double, float, false><<<kUpdBlocks, kFinBlock, 0, s>>>(
                nullptr, pc_next, r.as<float>(), q.as<float>(), n, st.dev, red, pq_blocks,
                red_upd);
        } else {
            ProfScope prof_u(kProfCgUpdateX64, n, 1, s);
            k_cg_update_pq<double, float><<<kUpdBlocks, kFinBlock, 0, s>>>(
                x, pc_next, r.as<float>(), q.as<float>(), n, st.dev, red, pq_blocks, red_upd);
        }
        HM_LAUNCH_CHECK();
        const TailSide side{x, pc_next, n, st.dev, 1};
        return pc_apply<float>(M, r.as<float>(), z.as<float>(), st.dev, red, kDotUpdateRz,
                               st.dev, n, s, x_beside_tail ? &side : nullptr);
    };
    // graph: each parity of the direction buffers' alternation gets its plain
    // iteration recorded once into a HIP graph (on the first iteration that
    // needs it) and replayed; replacement iterations stay direct launches.
    // Per-launch profiling records events, so it keeps direct launches.
    const bool use_graph = graph && !prof_on();
    hipGraph_t graphs[2] = {nullptr, nullptr};
    hipGraphExec_t execs[2] = {nullptr, nullptr};
    struct GraphGuard {
        hipGraph_t* g;
        hipGraphExec_t* e;
        ~GraphGuard() {
            for (int k = 0; k < 2; ++k) {
                if (e[k]) (void)hipGraphExecDestroy(e[k]);
                if (g[k]) (void)hipGraphDestroy(g[k]);
            }
        }
    } graph_guard{graphs, execs};
    float* const p_first = p.as<float>();
    if (maxiter <= 0) {  // no batch: read ||b|| here (b = 0 is solved by x = 0)
        HM_TRY(st.pull(s));
        stats->norm_b = stats->norm_r = st.host->norm_b;
        if (!(st.host->norm_b > 0.0)) {
            stats->converged = 1;
            return HM_OK;
        }
    }
    int kb = K;  // length of the next batch
    while (true) {
        const int k = std::min(kb, maxiter - enqueued);
        if (k <= 0) break;
        const int rep_at = rep_idx < k ? rep_idx : -1;
        for (int it = 0; it < k; ++it) {
            const bool replace = it == rep_at;
            if (use_graph && !replace) {
                const int par = p_cur == p_first ? 0 : 1;
                if (!execs[par]) {
                    HM_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
                    const int rc = plain_iteration(p_cur, p_next);
                    hipGraph_t g = nullptr;
                    const hipError_t e = hipStreamEndCapture(s, &g);
                    graphs[par] = g;
                    HM_TRY(rc);
                    HM_HIP(e);
                    HM_HIP(hipGraphInstantiate(&execs[par], g, nullptr, nullptr, 0));
                }
                HM_HIP(hipGraphLaunch(execs[par], s));
                std::swap(p_cur, p_next);
                continue;
            }
            HM_TRY(op_apply_pupdate<float>(A, z.as<float>(), p_cur, p_next, q.as<float>(), pq,
                                           st.dev, s));
            std::swap(p_cur, p_next);
            const bool x_in_tail = x_beside_tail && !replace;
            if (x_in_tail) {
                ProfScope prof_u(kProfCgUpdateR, n, 1, s);
                k_cg_update_pq<double, float, false><<<kUpdBlocks, kFinBlock, 0, s>>>(
                    nullptr, p_cur, r.as<float>(), q.as<float>(), n, st.dev, red, pq_blocks,
                    red_upd);
                HM_LAUNCH_CHECK();
            } else {
                ProfScope prof_u(kProfCgUpdateX64, n, 1, s);
                k_cg_update_pq<double, float><<<kUpdBlocks, kFinBlock, 0, s>>>(
                    x, p_cur, r.as<float>(), q.as<float>(), n, st.dev, red, pq_blocks, red_upd);
                HM_LAUNCH_CHECK();
            }
            if (replace) {
                // r = b - A x in float64, its norm for the convergence test.
                HM_TRY(op_residual<double>(A, b, x, r64.as<double>(), s));
                k_replace_r<<<kUpdBlocks, kFinBlock, 0, s>>>(r64.as<double>(), r.as<float>(), n, st.dev,
                                                  red_upd);
                HM_LAUNCH_CHECK();
            }
            const TailSide side{x, p_cur, n, st.dev, 1};
            HM_TRY(pc_apply<float>(M, r.as<float>(), z.as<float>(), st.dev, red, kDotUpdateRz,
                                   st.dev, n, s, x_in_tail ? &side : nullptr));
            if (replace) {
                k_note_rr<<<1, 1, 0, s>>>(st.dev);
                HM_LAUNCH_CHECK();
            }
        }
        const int rep_iter = rep_at >= 0 ? enqueued + rep_at + 1 : -1;
        const bool first_read = enqueued == 0;
        enqueued += k;
        HM_TRY(st.pull(s));
        if (first_read) {
            norm_b = st.host->norm_b;
            stats->norm_b = norm_b;
            stats->norm_r = norm_b;
            if (!(norm_b > 0.0)) {  // x = 0 solves b = 0 (the solve froze at once)
                stats->converged = 1;
                return HM_OK;
            }
            rr_replaced = rr_prev = norm_b * norm_b;
        }
        if (st.host->nonfinite) {
            set_error("non-finite value in the mixed solve");
            return HM_ERR_VALUE;
        }
        if (st.host->done) {
            if (st.host->iters == rep_iter) break;  // tested on a true residual
            // Confirm on the true residual. If the recursion drifted, restart
            // CG from the true residual (beta = 0): the old direction belongs
            // to a residual that may be far smaller than the true one.
            HM_TRY(op_residual<double>(A, b, x, r64.as<double>(), s));
            HM_TRY(launch_dot<double>(r64.as<double>(), r64.as<double>(), nullptr, n, st.dev,
                                      red, kDotAux, s));
            HM_TRY(st.pull(s));
            const double rr_true = st.host->aux;
            if (std::sqrt(rr_true) < rtol * norm_b) break;
            st.host->done = 0;
            st.host->rr = rr_true;
            st.host->beta = 0.0;
            HM_TRY(st.push(s));
            k_replace_r<<<kUpdBlocks, kFinBlock, 0, s>>>(r64.as<double>(), r.as<float>(), n, nullptr,
                                              red_upd);
            HM_LAUNCH_CHECK();
            HM_TRY(pc_apply<float>(M, r.as<float>(), z.as<float>(), st.dev, red, kDotRzFirst,
                                   nullptr, n, s));
            enqueued = st.host->iters;
            rr_replaced = rr_prev = rr_true;
            it_prev = enqueued;
            rep_idx = -1;
            continue;
        }
        const double rr = st.host->rr;
        const int its = st.host->iters;
        if (rep_at >= 0) rr_replaced = st.host->aux;
        if (its > it_prev && rr < rr_prev) rho2 = std::pow(rr / rr_prev, 1.0 / (its - it_prev));
        rr_prev = rr;
        it_prev = its;
        kb = next_batch(K, rr, rho2, rtol * rtol * norm_b * norm_b);
        // Iterations from now until ||r||^2 < delta^2 rr_replaced.
        const double target = delta * delta * rr_replaced;
        rep_idx = -1;
        int j_rep = -1;
        if (rr < target) {
            j_rep = 1;
        } else if (rho2 > 0.0 && rho2 < 1.0) {
            j_rep = (int)std::min<double>(
                1e9, std::max(1.0, std::ceil(std::log(target / rr) / std::log(rho2))));
        }
        if (j_rep > 0 && j_rep <= kb) {
            rep_idx = j_rep - 1;
            // A batch holds one replacement: it ends before the next one is
            // due, delta^2 further down at the measured rate.
            int gap = K;
            if (rho2 > 0.0 && rho2 < 1.0)
                gap = (int)std::max(1.0, std::ceil(std::log(delta * delta) / std::log(rho2)));
            kb = std::min(kb, j_rep + gap - 1);
        }
    }
    stats->iterations = st.host->iters;
    stats->norm_r = std::sqrt(st.host->rr);
    stats->converged = st.host->done;
    if (!st.host->done) {
        set_error("Conjugate gradient descent did not converge within %d iterations", maxiter);
        return HM_ERR_NOCONV;
    }
    return HM_OK;
}

// ------------------------------------------------------------ V-cycle set-up

void vc_free(hm_vc* vc) {
    for (size_t l = 0; l < vc->levels.size(); ++l) {
        VcLevel& L = vc->levels[l];
        if (L.owns_A && L.A) hm_dia_destroy(L.A);
        dev_free(L.b);
        dev_free(L.x);
        dev_free(L.t);
        dev_free(L.u);
    }
    dev_free(vc->dense_inv64);
    dev_free(vc->dense_inv32);
    dev_free(vc->tail_pack);
    dev_free(vc->tail_mat);
    free_reducer(&vc->red);
    delete vc;
}

int dia_alloc(int h, int w, hm_dia** out) {
    if (h < 1 || w < 1 || (long long)h * w > 0x7fffffffLL) {
        set_error("stored stencils need 1 <= h*w < 2^31 points, got %d x %d", h, w);
        return HM_ERR_ARG;
    }
    hm_dia* A = new hm_dia();
    A->h = h;
    A->w = w;
    A->n = (long long)h * w;
    int rc = dev_alloc(reinterpret_cast<void**>(&A->c64), sizeof(double) * kSymSlots * A->n);
    if (rc == HM_OK)
        rc = dev_alloc(reinterpret_cast<void**>(&A->c32), sizeof(float) * kSymSlots * A->n);
    if (rc == HM_OK) rc = dev_alloc(reinterpret_cast<void**>(&A->diag64), sizeof(double) * A->n);
    if (rc == HM_OK) rc = dev_alloc(reinterpret_cast<void**>(&A->diag32), sizeof(float) * A->n);
    if (rc == HM_OK) rc = dev_alloc(reinterpret_cast<void**>(&A->idiag64), sizeof(double) * A->n);
    if (rc == HM_OK) rc = dev_alloc(reinterpret_cast<void**>(&A->idiag32), sizeof(float) * A->n);
    if (rc != HM_OK) {
        hm_dia_destroy(A);
        return rc;
    }
    *out = A;
    return HM_OK;
}

int vc_create(hm_cf* cf, hm_dia* dia, int h, int w, int pre, int post, double omega, int direct,
              int prec, int max_levels, int coarse_sweeps, hipStream_t s, hm_vc** out) {
    if ((cf == nullptr) == (dia == nullptr)) {
        set_error("vcycle needs exactly one operator");
        return HM_ERR_ARG;
    }
    const long long n0 = cf ? cf->n : dia->n;
    if ((long long)h * w != n0) {
        if ((long long)h * w > direct) {
            set_error("dimension mismatch: shape (%d, %d) does not match an operator of size %lld",
                      h, w, n0);
            return HM_ERR_ARG;
        }
    }
    if (pre < 0 || post < 0) {
        set_error("num_pre_iter and num_post_iter must be non-negative");
        return HM_ERR_ARG;
    }
    if (max_levels < 0 || coarse_sweeps < 0) {
        set_error("max_levels and coarse_sweeps must be non-negative");
        return HM_ERR_ARG;
    }
    hm_vc* vc = new hm_vc();
    vc->prec = prec;
    vc->pre = pre;
    vc->post = post;
    vc->omega = omega;
    vc->direct_size = direct;
    const size_t es = prec == HM_F64 ? sizeof(double) : sizeof(float);
    int rc = HM_OK;
    hm_dia* fine = dia;
    hm_dia* tmp0 = nullptr;
    // A grid above the direct-solve size whose coarse grid (h/2, w/2) would
    // be empty ends the hierarchy as a smoothing-only level. The reference
    // recurses into the empty grid (vcycle.py:59-60 solves a 0 x 0 system),
    // so that level's cycle is its pre-smoothing from zero, a zero
    // correction and its post-smoothing: pre + post damped Jacobi sweeps
    // from zero.
    bool smooth_only = (long long)h * w > direct && (h / 2 == 0 || w / 2 == 0);
    // Level 0.
    VcLevel L0;
    L0.h = h;
    L0.w = w;
    L0.n = n0;
    if ((long long)h * w <= direct) {
        // Direct solve of the given operator (vcycle.py:59-60).
        if (cf) {
            rc = hm_dia_from_cf(cf, s, &tmp0);
            if (rc != HM_OK) { vc_free(vc); return rc; }
            L0.A = tmp0;
            L0.owns_A = true;
        } else {
            L0.A = dia;
        }
        L0.h = L0.A->h;
        L0.w = L0.A->w;
        if ((rc = dev_alloc(&L0.x, es * n0)) != HM_OK) { vc->levels.push_back(L0); vc_free(vc); return rc; }
        vc->levels.push_back(L0);
    } else {
        vc->cf = cf;
        L0.A = dia;  // null when matrix-free
        if (cf) rc = cf_prepare(cf, kCfDiag | (prec == HM_F32 ? kCfF32 : 0), s);
        if (rc == HM_OK && (pre >= 2 || post >= 2 || smooth_only)) {
            rc = dev_alloc(&L0.x, es * n0);
            if (rc == HM_OK) rc = dev_alloc(&L0.t, es * n0);
            if (rc == HM_OK) rc = dev_alloc(&L0.u, es * n0);
        }
        vc->levels.push_back(L0);
        if (rc != HM_OK) { vc_free(vc); return rc; }
        if (cf) {
            rc = hm_dia_from_cf(cf, s, &tmp0);  // exact level-0 coefficients for R A P
            if (rc != HM_OK) { vc_free(vc); return rc; }
            fine = tmp0;
        }
        int ch = h, cw = w;
        while (!smooth_only) {
            ch /= 2;
            cw /= 2;
            VcLevel Lc;
            Lc.h = ch;
            Lc.w = cw;
            Lc.n = (long long)ch * cw;
            rc = dia_alloc(ch, cw, &Lc.A);
            if (rc == HM_OK) {
                Lc.owns_A = true;
                rc = dia_galerkin(fine, Lc.A, s);  // also the float32 copies and diagonals
            }
            if (rc == HM_OK) rc = dev_alloc(&Lc.b, es * Lc.n);
            if (rc == HM_OK) rc = dev_alloc(&Lc.x, es * Lc.n);
            if (rc == HM_OK) rc = dev_alloc(&Lc.t, es * Lc.n);
            if (rc == HM_OK) rc = dev_alloc(&Lc.u, es * Lc.n);
            vc->levels.push_back(Lc);
            if (rc != HM_OK) break;
            fine = Lc.A;
            if (Lc.n <= direct) break;
            if (max_levels > 0 && (int)vc->levels.size() >= max_levels) break;
            if (ch / 2 == 0 || cw / 2 == 0) {
                smooth_only = true;
                break;
            }
        }
        if (tmp0) {
            HM_HIP(hipStreamSynchronize(s));
            hm_dia_destroy(tmp0);
            tmp0 = nullptr;
        }
        if (rc != HM_OK) { vc_free(vc); return rc; }
    }
    if (smooth_only) {
        if (pre + post == 0) {
            set_error("the (%d x %d) grid cannot be coarsened and V(0,0) smooths nothing on it; "
                      "give num_pre_iter + num_post_iter > 0",
                      vc->levels.back().h, vc->levels.back().w);
            vc_free(vc);
            return HM_ERR_ARG;
        }
        vc->coarse_sweeps = pre + post;
        rc = make_reducer(&vc->red, reducer_blocks(n0, h, w), 2);
        if (rc != HM_OK) { vc_free(vc); return rc; }
        HM_HIP(hipStreamSynchronize(s));
        *out = vc;
        return HM_OK;
    }
    const hm_dia* last = vc->levels.back().A;
    if (vc->levels.size() > 1 && last->n > kMaxDirect) {
        // Level cap reached above the dense-solve limit: approximate coarse solve.
        if (coarse_sweeps == 0) {
            set_error("the coarsest grid (%d x %d) exceeds the %d-point direct solve; give "
                      "coarse_sweeps > 0 or allow more levels", last->h, last->w, kMaxDirect);
            vc_free(vc);
            return HM_ERR_ARG;
        }
        vc->coarse_sweeps = coarse_sweeps;
        rc = make_reducer(&vc->red, reducer_blocks(n0, h, w), 2);
        if (rc != HM_OK) { vc_free(vc); return rc; }
        HM_HIP(hipStreamSynchronize(s));
        *out = vc;
        return HM_OK;
    }
    vc->n_direct = (int)last->n;
    rc = dev_alloc(reinterpret_cast<void**>(&vc->dense_inv64), sizeof(double) * last->n * last->n);
    if (rc == HM_OK)
        rc = dev_alloc(reinterpret_cast<void**>(&vc->dense_inv32), sizeof(float) * last->n * last->n);
    // The inverse's singularity flag is read with the set-up's closing
    // synchronisation rather than with one of its own.
    int* gj_status = nullptr;
    if (rc == HM_OK && last->n > 0)
        rc = dev_alloc(reinterpret_cast<void**>(&gj_status), sizeof(int));
    if (rc == HM_OK && last->n > 0)
        rc = dense_inverse(last, vc->dense_inv64, vc->dense_inv32, s, gj_status);
    if (rc == HM_OK) rc = make_reducer(&vc->red, reducer_blocks(n0, h, w), 2);
    if (rc == HM_OK && vc_tail_start(vc) < vc->levels.size() - 1)
        rc = prec == HM_F64 ? build_tail_pack<double>(vc, vc_tail_start(vc), s)
                            : build_tail_pack<float>(vc, vc_tail_start(vc), s);
#ifndef HM_TAIL_MATRIX
#define HM_TAIL_MATRIX 1
#endif
    if (HM_TAIL_MATRIX && rc == HM_OK && vc_tail_start(vc) < vc->levels.size() - 1)
        rc = prec == HM_F64 ? build_tail_matrix<double>(vc, vc_tail_start(vc), s)
                            : build_tail_matrix<float>(vc, vc_tail_start(vc), s);
    int singular = 0;
    if (rc == HM_OK && gj_status) {
        const hipError_t e =
            hipMemcpyAsync(&singular, gj_status, sizeof(int), hipMemcpyDeviceToHost, s);
        if (e != hipSuccess) {
            set_error("hipMemcpyAsync: %s", hipGetErrorString(e));
            rc = HM_ERR_HIP;
        }
    }
    const hipError_t es_ = hipStreamSynchronize(s);
    dev_free(gj_status);
    if (rc == HM_OK && es_ != hipSuccess) {
        set_error("hipStreamSynchronize: %s", hipGetErrorString(es_));
        rc = HM_ERR_HIP;
    }
    if (rc == HM_OK && singular) {
        set_error("coarsest-grid matrix is singular");
        rc = HM_ERR_VALUE;
    }
    if (rc != HM_OK) { vc_free(vc); return rc; }
    *out = vc;
    return HM_OK;
}

}  // namespace hm

// ===================================================================== C ABI

using namespace hm;

namespace hm {

// Mean duration of `reps` back-to-back launches of the finest grid's
// pre-smooth + residual + restriction (the V-cycle's first kernel, on the
// hierarchy's own operator and buffers, with r as the right-hand side) and
// of its prolongation + post-smooth, timed with HIP events on the stream.
template <typename T>
int time_fine_t(hm_vc* vc, const T* r, int reps, double* restrict_us, double* smooth_us,
                       hipStream_t s) {
    const VcLevel& lev = vc->levels[0];
    const VcLevel& cl = vc->levels[1];
    const LevelOp<T> op = vc_level_op<T>(vc, 0);
    const T omega = (T)vc->omega;
    const DotSink nodot{nullptr, Reducer{}, 0};
    DevBuf z;
    HM_TRY(z.alloc(sizeof(T) * lev.n));
    hipEvent_t e[3];
    for (auto& ev : e) HM_HIP(hipEventCreate(&ev));
    struct Ev { hipEvent_t* e; ~Ev() { for (int k = 0; k < 3; ++k) (void)hipEventDestroy(e[k]); } } g{e};
    const OutRestrict<T> rest{r, vptr<T>(cl.b), cl.w, cl.h, nullptr, nodot};
    const InJac0<T> jac{r, op.iD, omega};
    const InProlong<T, InJac0<T>> pro{jac, vptr<T>(cl.b), cl.w, cl.h};
    const OutSmooth<T> sm{r, op.iD, omega, z.as<T>(), nodot};
    for (int k = 0; k < 3; ++k) {  // warm-up
        HM_TRY(run_level<T>(op, jac, rest, nullptr, s));
        HM_TRY(run_level<T>(op, pro, sm, nullptr, s));
    }
    HM_HIP(hipEventRecord(e[0], s));
    for (int k = 0; k < reps; ++k) HM_TRY(run_level<T>(op, jac, rest, nullptr, s));
    HM_HIP(hipEventRecord(e[1], s));
    for (int k = 0; k < reps; ++k) HM_TRY(run_level<T>(op, pro, sm, nullptr, s));
    HM_HIP(hipEventRecord(e[2], s));
    HM_HIP(hipEventSynchronize(e[2]));
    float a = 0.f, b = 0.f;
    HM_HIP(hipEventElapsedTime(&a, e[0], e[1]));
    HM_HIP(hipEventElapsedTime(&b, e[1], e[2]));
    *restrict_us = 1e3 * a / reps;
    *smooth_us = 1e3 * b / reps;
    return HM_OK;
}

}  // namespace hm

extern "C" {

const char* hm_last_error(void) { return g_last_error.c_str(); }

int hm_version(void) { return 1; }

int hm_trimap_classes(const double* trimap, int64_t n, double fg_threshold, double bg_threshold,
                      uint8_t* cls, int64_t* counts, void* stream) {
    if (!trimap || !cls || !counts || n < 0) {
        set_error("hm_trimap_classes: null pointer or negative size");
        return HM_ERR_ARG;
    }
    return trimap_classes(trimap, n, fg_threshold, bg_threshold, cls, counts,
                          static_cast<hipStream_t>(stream));
}

int hm_cf_create(const double* image_hwc, int h, int w, double epsilon, int radius,
                 const uint8_t* cls, double lambda_value, int mode, void* stream, hm_cf** out) {
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (!image_hwc || !out || h < 1 || w < 1) {
        set_error("hm_cf_create: invalid image");
        return HM_ERR_ARG;
    }
    if (radius != 1) {
        set_error("only radius=1 (5x5 stencil) is implemented on the GPU, got radius=%d", radius);
        return HM_ERR_UNSUPPORTED;
    }
    if (mode < HM_OP_LAPLACIAN || mode > HM_OP_MASKED || (mode != HM_OP_LAPLACIAN && !cls)) {
        set_error("hm_cf_create: invalid mode or missing class map");
        return HM_ERR_ARG;
    }
    hm_cf* op = new hm_cf();
    op->h = h;
    op->w = w;
    op->n = (long long)h * w;
    op->eps = epsilon;
    op->lambda = lambda_value;
    op->mode = mode;
    op->skip_known = (mode == HM_OP_LAPLACIAN && cls) ? 1 : 0;
    const long long n = op->n;
    int rc = dev_alloc(reinterpret_cast<void**>(&op->I64), sizeof(double) * 3 * n);
    if (rc == HM_OK) rc = dev_alloc(reinterpret_cast<void**>(&op->I32), sizeof(float) * 3 * n);
    if (rc == HM_OK) rc = dev_alloc(reinterpret_cast<void**>(&op->wc), sizeof(double) * 10 * n);
    if (rc == HM_OK) rc = dev_alloc(reinterpret_cast<void**>(&op->wc32), sizeof(float) * 9 * n);
    if (rc == HM_OK) rc = dev_alloc(reinterpret_cast<void**>(&op->diag64), sizeof(double) * n);
    if (rc == HM_OK) rc = dev_alloc(reinterpret_cast<void**>(&op->diag32), sizeof(float) * n);
    if (rc == HM_OK) rc = dev_alloc(reinterpret_cast<void**>(&op->idiag64), sizeof(double) * n);
    if (rc == HM_OK) rc = dev_alloc(reinterpret_cast<void**>(&op->idiag32), sizeof(float) * n);
    if (rc == HM_OK && cls) {
        rc = dev_alloc(reinterpret_cast<void**>(&op->cls), n);
        if (rc == HM_OK) {
            hipError_t e = hipMemcpyAsync(op->cls, cls, n, hipMemcpyDeviceToDevice, s);
            if (e != hipSuccess) {
                set_error("hipMemcpyAsync: %s", hipGetErrorString(e));
                rc = HM_ERR_HIP;
            }
        }
    }
    if (rc == HM_OK) rc = cf_upload_image(image_hwc, op, s);
    if (rc == HM_OK) rc = cf_exact_setup(op, s);
    if (rc != HM_OK) {
        hm_cf_destroy(op);
        return rc;
    }
    *out = op;
    return HM_OK;
}

int hm_cf_destroy(hm_cf* op) {
    if (!op) return HM_OK;
    dev_free(op->I64);
    dev_free(op->I32);
    dev_free(op->cls);
    dev_free(op->wc);
    dev_free(op->wc32);
    dev_free(op->diag64);
    dev_free(op->diag32);
    dev_free(op->idiag64);
    dev_free(op->idiag32);
    delete op;
    return HM_OK;
}

int hm_cf_assemble_csr(hm_cf* op, double* values, int64_t* indices, int64_t* indptr,
                       void* stream) {
    if (!op || !values) {
        set_error("hm_cf_assemble_csr: null argument");
        return HM_ERR_ARG;
    }
    return cf_assemble(op, values, 25, 1, 0, indices, indptr, op->mode,
                       static_cast<hipStream_t>(stream));
}

int hm_cf_apply(hm_cf* op, int prec, const void* x, void* y, void* stream) {
    if (!op || !x || !y) {
        set_error("hm_cf_apply: null argument");
        return HM_ERR_ARG;
    }
    OpHandle A;
    A.cf = op;
    const DotSink nodot{nullptr, Reducer{}, 0};
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (prec == HM_F64)
        return op_apply<double>(A, static_cast<const double*>(x), static_cast<double*>(y), nodot,
                                nullptr, s);
    return op_apply<float>(A, static_cast<const float*>(x), static_cast<float*>(y), nodot, nullptr,
                           s);
}

int hm_cf_set_windows(hm_cf* op, int recompute) {
    if (!op || (recompute != 0 && recompute != 1)) {
        set_error("hm_cf_set_windows: invalid argument");
        return HM_ERR_ARG;
    }
    op->recompute = recompute;
    return HM_OK;
}

int hm_cf_diagonal(hm_cf* op, double* d, void* stream) {
    if (!op || !d) {
        set_error("hm_cf_diagonal: null argument");
        return HM_ERR_ARG;
    }
    HM_TRY(cf_prepare(op, kCfDiag, static_cast<hipStream_t>(stream)));
    HM_HIP(hipMemcpyAsync(d, op->diag64, sizeof(double) * op->n, hipMemcpyDeviceToDevice,
                          static_cast<hipStream_t>(stream)));
    return HM_OK;
}

int hm_cf_rhs(hm_cf* op, double* b, void* stream) {
    if (!op || !b || op->mode == HM_OP_LAPLACIAN) {
        set_error("hm_cf_rhs: needs a lambda or masked operator");
        return HM_ERR_ARG;
    }
    return cf_rhs(op, b, static_cast<hipStream_t>(stream));
}

int hm_dia_create(const double* coef, int h, int w, void* stream, hm_dia** out) {
    if (!coef || !out || h < 1 || w < 1) {
        set_error("hm_dia_create: invalid argument");
        return HM_ERR_ARG;
    }
    hipStream_t s = static_cast<hipStream_t>(stream);
    hm_dia* A = nullptr;
    HM_TRY(dia_alloc(h, w, &A));
    int rc = dia_pack(coef, A, s);
    if (rc == HM_OK) rc = dia_finish(A, s);
    if (rc != HM_OK) {
        hm_dia_destroy(A);
        return rc;
    }
    *out = A;
    return HM_OK;
}

int hm_dia_from_cf(hm_cf* op, void* stream, hm_dia** out) {
    hipStream_t s = static_cast<hipStream_t>(stream);
    hm_dia* A = nullptr;
    HM_TRY(dia_alloc(op->h, op->w, &A));
    // The 13 stored slots (dy, dx) >= (0, 0), assembled exactly in place.
    int rc = cf_assemble_sym(op, A, s);
    if (rc != HM_OK) {
        hm_dia_destroy(A);
        return rc;
    }
    *out = A;
    return HM_OK;
}

int hm_dia_destroy(hm_dia* A) {
    if (!A) return HM_OK;
    dev_free(A->c64);
    dev_free(A->c32);
    dev_free(A->diag64);
    dev_free(A->diag32);
    dev_free(A->idiag64);
    dev_free(A->idiag32);
    delete A;
    return HM_OK;
}

int hm_dia_apply(hm_dia* A, int prec, const void* x, void* y, void* stream) {
    if (!A || !x || !y) {
        set_error("hm_dia_apply: null argument");
        return HM_ERR_ARG;
    }
    OpHandle H;
    H.dia = A;
    const DotSink nodot{nullptr, Reducer{}, 0};
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (prec == HM_F64)
        return op_apply<double>(H, static_cast<const double*>(x), static_cast<double*>(y), nodot,
                                nullptr, s);
    return op_apply<float>(H, static_cast<const float*>(x), static_cast<float*>(y), nodot, nullptr,
                           s);
}

int hm_dia_coefficients(hm_dia* A, double* coef, void* stream) {
    if (!A || !coef) {
        set_error("hm_dia_coefficients: null argument");
        return HM_ERR_ARG;
    }
    return dia_unpack(A, coef, static_cast<hipStream_t>(stream));
}

int hm_csr_create(int64_t n, int64_t nnz, const int64_t* indptr, const int64_t* indices,
                  const double* values, hm_csr** out) {
    if (n < 0 || nnz < 0 || !indptr || !out) {
        set_error("hm_csr_create: invalid argument");
        return HM_ERR_ARG;
    }
    hm_csr* A = new hm_csr();
    A->n = n;
    A->nnz = nnz;
    int rc = dev_alloc(reinterpret_cast<void**>(&A->indptr), sizeof(int64_t) * (n + 1));
    if (rc == HM_OK) rc = dev_alloc(reinterpret_cast<void**>(&A->indices), sizeof(int64_t) * nnz);
    if (rc == HM_OK) rc = dev_alloc(reinterpret_cast<void**>(&A->values), sizeof(double) * nnz);
    if (rc == HM_OK) {
        if (hipMemcpy(A->indptr, indptr, sizeof(int64_t) * (n + 1), hipMemcpyDeviceToDevice) != hipSuccess ||
            (nnz > 0 && hipMemcpy(A->indices, indices, sizeof(int64_t) * nnz, hipMemcpyDeviceToDevice) != hipSuccess) ||
            (nnz > 0 && hipMemcpy(A->values, values, sizeof(double) * nnz, hipMemcpyDeviceToDevice) != hipSuccess)) {
            set_error("hm_csr_create: copy failed");
            rc = HM_ERR_HIP;
        }
    }
    if (rc != HM_OK) {
        hm_csr_destroy(A);
        return rc;
    }
    *out = A;
    return HM_OK;
}

int hm_csr_destroy(hm_csr* A) {
    if (!A) return HM_OK;
    dev_free(A->indptr);
    dev_free(A->indices);
    dev_free(A->values);
    delete A;
    return HM_OK;
}

int hm_csr_apply(hm_csr* A, const double* x, double* y, void* stream) {
    return csr_apply(A, x, y, nullptr, nullptr, 0, static_cast<hipStream_t>(stream));
}

int hm_vcycle_create(hm_cf* cf, hm_dia* dia, int h, int w, int num_pre_iter, int num_post_iter,
                     double omega, int direct_solve_size, int prec, void* stream, hm_vc** out) {
    if (!out) {
        set_error("hm_vcycle_create: null output");
        return HM_ERR_ARG;
    }
    return vc_create(cf, dia, h, w, num_pre_iter, num_post_iter, omega, direct_solve_size, prec,
                     0, 0, static_cast<hipStream_t>(stream), out);
}

int hm_vcycle_create_ex(hm_cf* cf, hm_dia* dia, int h, int w, int num_pre_iter,
                        int num_post_iter, double omega, int direct_solve_size, int prec,
                        int max_levels, int coarse_sweeps, void* stream, hm_vc** out) {
    if (!out) {
        set_error("hm_vcycle_create_ex: null output");
        return HM_ERR_ARG;
    }
    return vc_create(cf, dia, h, w, num_pre_iter, num_post_iter, omega, direct_solve_size, prec,
                     max_levels, coarse_sweeps, static_cast<hipStream_t>(stream), out);
}

int hm_vcycle_create_dense(const double* A_dense, int n, int prec, void* stream, hm_vc** out) {
    if (!A_dense || !out || n < 1) {
        set_error("hm_vcycle_create_dense: invalid argument");
        return HM_ERR_ARG;
    }
    hipStream_t s = static_cast<hipStream_t>(stream);
    hm_vc* vc = new hm_vc();
    vc->prec = prec;
    VcLevel L0;
    L0.h = 1;
    L0.w = n;
    L0.n = n;
    vc->levels.push_back(L0);
    vc->n_direct = n;
    int rc = dev_alloc(reinterpret_cast<void**>(&vc->dense_inv64), sizeof(double) * (size_t)n * n);
    if (rc == HM_OK)
        rc = dev_alloc(reinterpret_cast<void**>(&vc->dense_inv32), sizeof(float) * (size_t)n * n);
    if (rc == HM_OK) rc = dense_inverse_from_dense(A_dense, n, vc->dense_inv64, vc->dense_inv32, s);
    if (rc == HM_OK) rc = make_reducer(&vc->red, reducer_blocks(n, 1, n), 2);
    if (rc != HM_OK) {
        vc_free(vc);
        return rc;
    }
    *out = vc;
    return HM_OK;
}

int hm_vcycle_destroy(hm_vc* vc) {
    if (vc) vc_free(vc);
    return HM_OK;
}

int hm_vcycle_levels(hm_vc* vc) { return vc ? (int)vc->levels.size() : 0; }

int hm_vcycle_set_launches(hm_vc* vc, int pairs) {
    if (!vc) {
        hm::set_error("null V-cycle");
        return HM_ERR_ARG;
    }
    if (pairs < 0 || pairs > (hm::kVcChainUp | hm::kVcPairDown | hm::kVcTailMatrix)) {
        hm::set_error("launch structure must be 0..7, got %d", pairs);
        return HM_ERR_ARG;
    }
    vc->pairs = pairs;
    return HM_OK;
}

int hm_vcycle_apply(hm_vc* vc, const void* r, void* z, void* stream) {
    if (!vc || !r || !z) {
        set_error("hm_vcycle_apply: null argument");
        return HM_ERR_ARG;
    }
    return vcycle_apply(vc, r, z, nullptr, nullptr, 0, nullptr, static_cast<hipStream_t>(stream));
}

int hm_vcycle_time_fine(hm_vc* vc, const void* r, int reps, double* restrict_us,
                        double* smooth_us, void* stream) {
    if (!vc || !r || reps < 1 || !restrict_us || !smooth_us) {
        set_error("hm_vcycle_time_fine: invalid argument");
        return HM_ERR_ARG;
    }
    if (vc->levels.size() < 2 || vc->pre != 1 || vc->post != 1) {
        set_error("hm_vcycle_time_fine: needs a V(1,1) cycle of at least two grids");
        return HM_ERR_ARG;
    }
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (vc->prec == HM_F64)
        return hm::time_fine_t<double>(vc, static_cast<const double*>(r), reps, restrict_us,
                                       smooth_us, s);
    return hm::time_fine_t<float>(vc, static_cast<const float*>(r), reps, restrict_us,
                                  smooth_us, s);
}

int hm_copy(void* dst, const void* src, int64_t bytes, int kind, void* stream) {
    if (bytes < 0 || (bytes > 0 && (!dst || !src))) {
        set_error("hm_copy: invalid argument");
        return HM_ERR_ARG;
    }
    if (bytes == 0) return HM_OK;
    hipStream_t s = static_cast<hipStream_t>(stream);
    const hipMemcpyKind k = kind == 0 ? hipMemcpyDeviceToHost
                                      : (kind == 1 ? hipMemcpyHostToDevice : hipMemcpyDeviceToDevice);
    HM_HIP(hipMemcpyAsync(dst, src, (size_t)bytes, k, s));
    if (kind != 2) HM_HIP(hipStreamSynchronize(s));
    return HM_OK;
}

int hm_cg_solve(const hm_cg_args* a, hm_cg_stats* stats, void* stream) {
    if (!a || !stats) {
        set_error("hm_cg_solve: null argument");
        return HM_ERR_ARG;
    }
    const int nops = (a->cf != nullptr) + (a->dia != nullptr) + (a->csr != nullptr) +
                     (a->lkm != nullptr) + (a->a_cb != nullptr);
    if (nops != 1 || !a->b || !a->x || (a->a_cb && a->n < 0)) {
        set_error("hm_cg_solve: needs exactly one operator and b, x");
        return HM_ERR_ARG;
    }
    const int k = a->pc_kind;
    if (k < HM_PC_NONE || k > HM_PC_ICHOL || (k == HM_PC_JACOBI && !a->inv_diag) ||
        (k == HM_PC_VCYCLE && !a->vc) || (k == HM_PC_CSR && !a->pc_csr) ||
        (k == HM_PC_HOST && !a->m_cb) || (k == HM_PC_ICHOL && !a->ichol)) {
        set_error("hm_cg_solve: invalid preconditioner");
        return HM_ERR_ARG;
    }
    if (k == HM_PC_VCYCLE && a->vc->prec != HM_F64) {
        set_error("hm_cg_solve: the V-cycle must be float64 for a float64 solve");
        return HM_ERR_ARG;
    }
    OpHandle A;
    A.cf = a->cf;
    A.dia = a->dia;
    A.csr = a->csr;
    A.lkm = a->lkm;
    A.cb = a->a_cb;
    A.cb_ctx = a->a_ctx;
    A.cb_n = a->n;
    if (k == HM_PC_VCYCLE && a->vc->levels[0].n != A.n()) {
        set_error("dimension mismatch: V-cycle of size %lld for an operator of size %lld",
                  a->vc->levels[0].n, A.n());
        return HM_ERR_ARG;
    }
    if (k == HM_PC_ICHOL && ichol_rows(a->ichol) != A.n()) {
        set_error("dimension mismatch: ichol factor of size %lld for an operator of size %lld",
                  ichol_rows(a->ichol), A.n());
        return HM_ERR_ARG;
    }
    PcHandle M;
    M.kind = k;
    M.inv_diag = a->inv_diag;
    M.vc = a->vc;
    M.csr = a->pc_csr;
    M.ichol = a->ichol;
    M.cb = a->m_cb;
    M.cb_ctx = a->m_ctx;
    PcgOptions opt;
    opt.atol = a->atol;
    opt.rtol = a->rtol;
    opt.maxiter = a->maxiter;
    opt.reorth = a->reorthogonalize;
    opt.check_every = a->check_every;
    opt.callback = a->callback;
    opt.cb_ctx = a->cb_ctx;
    return pcg_t<double>(A, M, a->b, a->x0, a->x, opt, stats, static_cast<hipStream_t>(stream));
}

int hm_cg_solve_reliable(hm_cf* cf, hm_dia* dia, hm_vc* vc32, const double* b, double* x,
                         double rtol, double delta, int maxiter, int check_every, int graph,
                         hm_cg_stats* stats, void* stream) {
    if ((cf == nullptr) == (dia == nullptr) || !vc32 || !b || !x || !stats ||
        vc32->prec != HM_F32 || !(delta > 0.0 && delta < 1.0) || (graph != 0 && graph != 1)) {
        set_error("hm_cg_solve_reliable: needs one operator, a float32 V-cycle, 0 < delta < 1 "
                  "and graph 0 or 1");
        return HM_ERR_ARG;
    }
    hipStream_t user = static_cast<hipStream_t>(stream);
    if (!graph) return pcg_reliable(cf, dia, vc32, b, x, rtol, delta, maxiter, check_every,
                                    false, stats, user);
    // Graphs cannot be captured on the legacy default stream: the solve runs
    // on a stream of this thread's own, ordered after and before the caller's.
    static thread_local hipStream_t own = nullptr;
    static thread_local hipEvent_t ev_in = nullptr, ev_out = nullptr;
    if (!own) {
        HM_HIP(hipStreamCreateWithFlags(&own, hipStreamNonBlocking));
        HM_HIP(hipEventCreateWithFlags(&ev_in, hipEventDisableTiming));
        HM_HIP(hipEventCreateWithFlags(&ev_out, hipEventDisableTiming));
    }
    HM_HIP(hipEventRecord(ev_in, user));
    HM_HIP(hipStreamWaitEvent(own, ev_in, 0));
    const int rc = pcg_reliable(cf, dia, vc32, b, x, rtol, delta, maxiter, check_every, true,
                                stats, own);
    (void)hipStreamSynchronize(own);
    if (hipEventRecord(ev_out, own) == hipSuccess) (void)hipStreamWaitEvent(user, ev_out, 0);
    return rc;
}

int hm_cg_solve_mixed(hm_cf* cf, hm_dia* dia, hm_vc* vc32, const double* b, double* x,
                      double rtol, double inner_rtol, int max_outer, int maxiter,
                      hm_cg_stats* stats, void* stream) {
    if ((cf == nullptr) == (dia == nullptr) || !vc32 || !b || !x || !stats ||
        vc32->prec != HM_F32) {
        set_error("hm_cg_solve_mixed: needs one operator and a float32 V-cycle");
        return HM_ERR_ARG;
    }
    return pcg_mixed(cf, dia, vc32, b, x, rtol, inner_rtol, max_outer, maxiter, stats,
                     static_cast<hipStream_t>(stream));
}

}  // extern "C"

extern "C" {

int hm_cf_system_indptr(hm_cf* op, int64_t* indptr, int64_t* nnz, void* stream) {
    if (!op || !indptr || !nnz) {
        set_error("hm_cf_system_indptr: null argument");
        return HM_ERR_ARG;
    }
    hipStream_t s = static_cast<hipStream_t>(stream);
    DevBuf counts;
    HM_TRY(counts.alloc(sizeof(int64_t) * op->n));
    HM_TRY(cf_system_counts(op, counts.as<int64_t>(), s));
    return exclusive_offsets(counts.as<int64_t>(), indptr, op->n, nnz, s);
}

int hm_cf_system_fill(hm_cf* op, const int64_t* indptr, int64_t* indices, double* values,
                      void* stream) {
    if (!op || !indptr || !indices || !values) {
        set_error("hm_cf_system_fill: null argument");
        return HM_ERR_ARG;
    }
    return cf_system_fill(op, indptr, indices, values, static_cast<hipStream_t>(stream));
}

}  // extern "C"

namespace hm {

__global__ void k_diag_apply(const double* __restrict__ inv_diag, const double* __restrict__ r,
                             double* __restrict__ z, long long n) {
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) z[i] = r[i] * inv_diag[i];
}

// alpha = clip(x on unknown pixels, trimap on known pixels, 0, 1)
// (estimate_alpha_cf.py:72-76); lambda path: clip(x, 0, 1).
__global__ void k_alpha_finalize(const double* __restrict__ x, const double* __restrict__ trimap,
                                 const uint8_t* __restrict__ cls, double* __restrict__ alpha,
                                 long long n) {
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double v = (cls && cls[i] != 0) ? trimap[i] : x[i];
    v = v < 0.0 ? 0.0 : v;
    v = v > 1.0 ? 1.0 : v;
    alpha[i] = v;
}

}  // namespace hm

extern "C" {

int hm_diag_apply(const double* inv_diag, const double* r, double* z, int64_t n, void* stream) {
    if (!inv_diag || !r || !z || n < 0) {
        set_error("hm_diag_apply: invalid argument");
        return HM_ERR_ARG;
    }
    if (n == 0) return HM_OK;
    k_diag_apply<<<div_up(n, 256), 256, 0, static_cast<hipStream_t>(stream)>>>(inv_diag, r, z, n);
    HM_LAUNCH_CHECK();
    return HM_OK;
}

int hm_alpha_finalize(const double* x, const double* trimap, const uint8_t* cls, double* alpha,
                      int64_t n, void* stream) {
    if (!x || !alpha || n < 0 || (cls && !trimap)) {
        set_error("hm_alpha_finalize: invalid argument");
        return HM_ERR_ARG;
    }
    if (n == 0) return HM_OK;
    k_alpha_finalize<<<div_up(n, 256), 256, 0, static_cast<hipStream_t>(stream)>>>(x, trimap, cls,
                                                                                   alpha, n);
    HM_LAUNCH_CHECK();
    return HM_OK;
}

}  // extern "C"

// ----------------------------------------------------------------- profiling

namespace hm {
namespace {
struct ProfRecord {
    hipEvent_t start, stop;
    int kind;
    int prec;
    long long points;
};
bool g_prof_on = false;
std::vector<ProfRecord> g_prof;
std::vector<std::pair<hipEvent_t, hipEvent_t>> g_prof_pool;
size_t g_prof_used = 0;
}  // namespace

bool prof_on() { return g_prof_on; }

ProfScope::ProfScope(int kind, long long points, int prec, hipStream_t s) : slot(-1), stream(s) {
    if (!g_prof_on) return;
    if (g_prof_used == g_prof_pool.size()) {
        hipEvent_t a, b;
        if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) return;
        g_prof_pool.emplace_back(a, b);
    }
    ProfRecord r{g_prof_pool[g_prof_used].first, g_prof_pool[g_prof_used].second, kind, prec,
                 points};
    ++g_prof_used;
    if (hipEventRecord(r.start, s) != hipSuccess) return;
    slot = (int)g_prof.size();
    g_prof.push_back(r);
}

ProfScope::~ProfScope() {
    if (slot >= 0) (void)hipEventRecord(g_prof[slot].stop, stream);
}

}  // namespace hm

extern "C" {

int hm_prof_set(int enabled) {
    hm::g_prof_on = enabled != 0;
    hm::g_prof.clear();
    hm::g_prof_used = 0;
    return HM_OK;
}

int hm_prof_query(int kind, int prec, double* total_ms, int64_t* launches, int64_t* points) {
    if (!total_ms || !launches || !points) {
        set_error("hm_prof_query: null argument");
        return HM_ERR_ARG;
    }
    double ms = 0.0;
    int64_t cnt = 0, pts = 0;
    for (const auto& r : hm::g_prof) {
        if (r.kind != kind || r.prec != prec) continue;
        HM_HIP(hipEventSynchronize(r.stop));
        float e = 0.f;
        HM_HIP(hipEventElapsedTime(&e, r.start, r.stop));
        ms += e;
        ++cnt;
        pts += r.points;
    }
    *total_ms = ms;
    *launches = cnt;
    *points = pts;
    return HM_OK;
}

}  // extern "C"
