// Level operators, vector kernels and the V-cycle building blocks shared by
// the single-system drivers (solver.hip) and the batched drivers (batch.hip).
#pragma once

#include <type_traits>

#include <algorithm>
#include <vector>

#include "common.h"
#include "engine.h"
#include "internal.h"

namespace hm {

// ------------------------------------------------------------- launch helpers

// level: the grid's place in the hierarchy (0 finest); profiling classes
// and kernel names separate the finest grid, the first Galerkin grid and the
// coarser ones.
template <typename T, class Op, class In, class Out>
int launch_tile(const Op& op, const In& in, const Out& out, int h, int w, int h2, int w2,
                const CgState* guard, hipStream_t s, int level = 0) {
    const TileMap map = tile_map(out, h, w, h2, w2);
    if (out.dot.st != nullptr && (long long)map.ntiles > out.dot.red.max_blocks) {
        set_error("reduction buffer too small for %u blocks", map.ntiles);
        return HM_ERR_ARG;
    }
    const int base = level == 0 ? (int)Op::kProfBase
                                : (level == 1 ? (int)kProfLevel1Apply : (int)kProfCoarseApply);
    ProfScope prof(base + Out::kProf, (long long)h * w, sizeof(T) == 4, s);
    if constexpr (std::is_same<Op, CfOp<T>>::value || std::is_same<Op, CfOp<T, true>>::value) {
        tile_kernel<T, Op, In, Out, 0><<<map.ntiles, kBlock, 0, s>>>(op, in, out, h, w, guard, map);
    } else {
        if (level == 0)
            tile_kernel<T, Op, In, Out, 0><<<map.ntiles, kBlock, 0, s>>>(op, in, out, h, w, guard,
                                                                         map);
        else if (level == 1)
            tile_kernel<T, Op, In, Out, 1><<<map.ntiles, kBlock, 0, s>>>(op, in, out, h, w, guard,
                                                                         map);
        else
            tile_kernel<T, Op, In, Out, 2><<<map.ntiles, kBlock, 0, s>>>(op, in, out, h, w, guard,
                                                                         map);
    }
    HM_LAUNCH_CHECK();
    if (out.dot.st != nullptr && !out.dot.defer)
        return launch_finalize(1, out.dot.red, map.ntiles, out.dot.st, out.dot.kind, guard, s);
    return HM_OK;
}

template <typename T>
struct LevelOp {
    bool is_cf = false;
    bool cf_rec = false;  // float32 matrix-free passes recompute the windows
    int level = 0;  // place in the hierarchy (profiling classes and kernel names only)
    CfOp<T> cf{};
    DiaOp<T> dia{};
    const T* iD = nullptr;  // 1 / diag
    int h = 0, w = 0;
};

template <typename T>
const T* pick(const double* a, const float* b);
template <>
inline const double* pick<double>(const double* a, const float*) { return a; }
template <>
inline const float* pick<float>(const double*, const float* b) { return b; }

template <typename T>
LevelOp<T> cf_level(const hm_cf* op, int force_mode = -1) {
    LevelOp<T> L;
    L.is_cf = true;
    L.cf.d = CfDev{op->h, op->w, op->mode, op->skip_known, op->eps, op->lambda, op->cls};
    L.cf.I = pick<T>(op->I64, op->I32);
    L.cf.W = pick<T>(op->wc, op->wc32);
    L.cf.force_mode = force_mode;
    L.cf_rec = sizeof(T) == 4 && op->recompute;
    L.iD = pick<T>(op->idiag64, op->idiag32);
    L.h = op->h;
    L.w = op->w;
    return L;
}

template <typename T>
LevelOp<T> dia_level(const hm_dia* A) {
    LevelOp<T> L;
    L.is_cf = false;
    L.dia = DiaOp<T>{A->h, A->w, pick<T>(A->c64, A->c32)};
    L.iD = pick<T>(A->idiag64, A->idiag32);
    L.h = A->h;
    L.w = A->w;
    return L;
}

template <typename T, class In, class Out>
int run_level(const LevelOp<T>& L, const In& in, const Out& out, const CgState* guard,
              hipStream_t s) {
    if (L.is_cf) {
        if constexpr (sizeof(T) == 4) {
            if (L.cf_rec) {
                const CfOp<T, true> rec{L.cf.d, L.cf.I, nullptr, L.cf.force_mode};
                return launch_tile<T>(rec, in, out, L.h, L.w, L.h / 2, L.w / 2, guard, s);
            }
        }
        return launch_tile<T>(L.cf, in, out, L.h, L.w, L.h / 2, L.w / 2, guard, s);
    }
    // Coarse grids below kSmallGrid points are latency-bound: one point per
    // thread with its coefficients loaded before the tile barrier. Larger
    // grids keep the bandwidth-oriented shape of the finest level.
#ifndef HM_SMALL_GRID_LOG2
#define HM_SMALL_GRID_LOG2 18
#endif
    constexpr long long kSmallGrid = 1LL << HM_SMALL_GRID_LOG2;
    if (L.level > 0 && (long long)L.h * L.w < kSmallGrid) {
        const DiaOp<T, true> op{L.dia.h, L.dia.w, L.dia.coef};
        return launch_tile<T>(op, in, out.small(), L.h, L.w, L.h / 2, L.w / 2, guard, s,
                              L.level);
    }
    return launch_tile<T>(L.dia, in, out, L.h, L.w, L.h / 2, L.w / 2, guard, s, L.level);
}

// ------------------------------------------------------------- vector kernels

constexpr int kVecPerThread = 4;

inline unsigned vec_blocks(long long n) { return div_up(n, (long long)kBlock * kVecPerThread); }

// x += alpha p; r -= alpha q; ||r||^2 and the convergence test (cg.py:90-100).
template <typename T>
__global__ __launch_bounds__(kBlock) void k_cg_update(T* __restrict__ x, const T* __restrict__ p,
                                                      T* __restrict__ r, const T* __restrict__ q,
                                                      long long n, CgState* st, Reducer red) {
    if (solve_done(st)) return;
    const T alpha = (T)st->alpha;
    double acc = 0.0;
    const long long base = (long long)blockIdx.x * kBlock * kVecPerThread + threadIdx.x;
#pragma unroll
    for (int k = 0; k < kVecPerThread; ++k) {
        const long long i = base + (long long)k * kBlock;
        if (i < n) {
            x[i] += alpha * p[i];
            const T ri = r[i] - alpha * q[i];
            r[i] = ri;
            acc += (double)ri * (double)ri;
        }
    }
    double v[1] = {acc};
    block_partials<kBlock, 1>(v, red);
}

// p = z + beta p (cg.py:112-113).
template <typename T>
__global__ __launch_bounds__(kBlock) void k_cg_pupdate(T* __restrict__ p, const T* __restrict__ z,
                                                       long long n, const CgState* st) {
    if (solve_done(st)) return;
    const T beta = (T)st->beta;
    const long long base = (long long)blockIdx.x * kBlock * kVecPerThread + threadIdx.x;
#pragma unroll
    for (int k = 0; k < kVecPerThread; ++k) {
        const long long i = base + (long long)k * kBlock;
        if (i < n) p[i] = p[i] * beta + z[i];
    }
}

// z = r * inv_diag (jacobi.py:24-31), or z = r without a preconditioner; <r, z>.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_diag_pc(const T* __restrict__ r,
                                                    const double* __restrict__ inv_diag,
                                                    T* __restrict__ z, long long n, CgState* st,
                                                    Reducer red, int kind) {
    if (solve_done(st)) return;
    double acc = 0.0;
    const long long base = (long long)blockIdx.x * kBlock * kVecPerThread + threadIdx.x;
#pragma unroll
    for (int k = 0; k < kVecPerThread; ++k) {
        const long long i = base + (long long)k * kBlock;
        if (i < n) {
            const T zi = inv_diag ? (T)((double)r[i] * inv_diag[i]) : r[i];
            z[i] = zi;
            acc += (double)r[i] * (double)zi;
        }
    }
    if (st != nullptr) {
        double v[1] = {acc};
        block_partials<kBlock, 1>(v, red);
    }
}

// <a, b> (and <a - c, b> when c is given, for the Polak-Ribiere beta of
// cg.py:104-106) into the state.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_dot(const T* __restrict__ a, const T* __restrict__ b,
                                                const T* __restrict__ c, long long n, CgState* st,
                                                Reducer red, int kind) {
    if (kind != kDotAux && kind < kDotNormB && solve_done(st)) return;
    double acc[2] = {0.0, 0.0};
    const long long base = (long long)blockIdx.x * kBlock * kVecPerThread + threadIdx.x;
#pragma unroll
    for (int k = 0; k < kVecPerThread; ++k) {
        const long long i = base + (long long)k * kBlock;
        if (i < n) {
            acc[0] += (double)a[i] * (double)b[i];
            if (c) acc[1] += ((double)a[i] - (double)c[i]) * (double)b[i];
        }
    }
    block_partials<kBlock, 2>(acc, red);
}

// <a, b> (or the Polak-Ribiere pair when c is given) into the solve state.
template <typename T>
int launch_dot(const T* a, const T* b, const T* c, long long n, CgState* st, const Reducer& red,
               int kind, hipStream_t s) {
    const unsigned g = vec_blocks(n);
    k_dot<T><<<g, kBlock, 0, s>>>(a, b, c, n, st, red, kind);
    HM_LAUNCH_CHECK();
    return launch_finalize(c ? 2 : 1, red, g, st, c ? kDotReorth : kind,
                           kind == kDotAux || kind >= kDotNormB ? nullptr : st, s);
}

template <typename T>
__global__ void k_copy(const T* __restrict__ a, T* __restrict__ b, long long n,
                       const CgState* st) {
    if (solve_done(st)) return;
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = a[i];
}

template <typename T>
__global__ void k_fill(T* __restrict__ a, T v, long long n) {
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = v;
}

// out = base + P e with base = omega b / D (pre = 1), a stored vector, or 0.
template <typename T, class Base>
__global__ void k_prolong_add(Base base, InProlong<T, Base> pro, T* __restrict__ out, int h,
                              int w, const CgState* st) {
    if (solve_done(st)) return;
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long long)h * w) return;
    const int y = (int)(i / w), x = (int)(i - (long long)y * w);
    out[i] = pro.load(x, y, i);
}

// x = inv(A_c) b on the coarsest grid (n <= 1024), optional <b, x>.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_dense_apply(const T* __restrict__ inv,
                                                        const T* __restrict__ b,
                                                        T* __restrict__ x, int n, CgState* st,
                                                        Reducer red, const CgState* guard) {
    if (solve_done(guard)) return;
    double acc = 0.0;
    for (int i = threadIdx.x + blockIdx.x * kBlock; i < n; i += kBlock * gridDim.x) {
        T s = 0;
#pragma unroll 8
        for (int j = 0; j < n; ++j) s += inv[(long long)i * n + j] * b[j];
        x[i] = s;
        acc += (double)b[i] * (double)s;
    }
    if (st != nullptr) {  // a single workgroup: one partial, finalised by the host's next launch
        double v[1] = {acc};
        block_partials<kBlock, 1>(v, red);
    }
}

// First damped-Jacobi iterate from zero: x = omega b / D (vcycle.py:35-38),
// with iD = 1 / D.
template <typename T>
__global__ void k_jac0(const T* __restrict__ b, const T* __restrict__ iD, T omega,
                       T* __restrict__ x, long long n, const CgState* guard) {
    if (solve_done(guard)) return;
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] = vc_jac0(omega, b[i], iD[i]);
}

// ------------------------------------------------------------------ V-cycle

template <typename T>
T* vptr(void* p) { return static_cast<T*>(p); }

template <typename T>
LevelOp<T> vc_level_op(const hm_vc* vc, size_t l) {
    if (l == 0 && vc->cf) return cf_level<T>(vc->cf);
    LevelOp<T> L = dia_level<T>(vc->levels[l].A);
    L.level = (int)l;
    return L;
}

template <typename T>
const T* vc_inv(const hm_vc* vc) { return pick<T>(vc->dense_inv64, vc->dense_inv32); }

template <typename T, class Base>
int post_stage(const LevelOp<T>& op, const Base& base, const T* b, const T* e, int h2, int w2,
               int post, T omega, T* target, T* other, DotSink fin, const T* r_top,
               const CgState* guard, hipStream_t s, const Reducer& red) {
    DotSink nodot{nullptr, Reducer{}, 0};
    InProlong<T, Base> pro{base, e, w2, h2};
    if (post == 0) {
        k_prolong_add<T, Base><<<div_up((long long)op.h * op.w, 256), 256, 0, s>>>(
            base, pro, target, op.h, op.w, guard);
        HM_LAUNCH_CHECK();
        if (fin.st != nullptr) {
            const long long n = (long long)op.h * op.w;
            HM_TRY(launch_dot<T>(r_top, target, nullptr, n, fin.st, red, fin.kind, s));
        }
        return HM_OK;
    }
    const T* prev = nullptr;
    for (int k = 0; k < post; ++k) {
        T* outp = ((post - 1 - k) % 2 == 0) ? target : other;
        const DotSink ds = (k == post - 1) ? fin : nodot;
        OutSmooth<T> sm{b, op.iD, omega, outp, ds};
        if (k == 0)
            HM_TRY(run_level<T>(op, pro, sm, guard, s));
        else
            HM_TRY(run_level<T>(op, InVec<T>{prev}, sm, guard, s));
        prev = outp;
    }
    return HM_OK;
}

// ------------------------------------------------------------- V-cycle tail
//
// The coarsest grids are processed by one 1024-thread workgroup in a single
// launch. Every vector of every tail grid lives in LDS with a zero halo of two
// points, so the 25-point stencil, the restriction and the prolongation run
// without bounds tests (out-of-grid neighbours read zeros, which is the
// reference's dropped taps). The stencils of all but possibly the top tail
// grid are staged in LDS too; the top grid's may stay in global memory (L2)
// when they do not fit, which lets a larger grid join the tail. Only the top
// grid's right-hand side is read from and its correction written to global
// memory. On these grids each separate launch would cost far more than its
// arithmetic.

constexpr int kTailMaxLevels = 24;
constexpr int kTailBlock = 1024;
constexpr int kTailLdsBytes = 152 * 1024;
#ifndef HM_TAIL_MAX_TOP
#define HM_TAIL_MAX_TOP 1024
#endif
constexpr int kTailMaxTop = HM_TAIL_MAX_TOP;  // largest top grid of a tail (points)

template <typename T>
struct TailLevel {
    int h, w;
    const T* coef;  // [13][n] symmetric storage (global)
    T* b;           // right-hand side (global; only the top level's is read)
    T* x;           // result (global; only the top level's is written)
    int off;        // LDS offset of this level's vectors (in T units)
    int coef_lds;   // 1: coefficients staged in LDS at coef_off
    int coef_off;
};

template <typename T>
struct TailParams {
    TailLevel<T> lv[kTailMaxLevels];
    int nlev;        // tail grids including the direct one (last)
    int pre, post;
    T omega;
    const T* inv;    // dense inverse of the last grid (global)
    int inv_off;     // LDS offset of the staged inverse, or -1 (read from global)
    const T* pack;   // constant LDS contents [pack_off, pack_off + pack_count) (global)
    int pack_off, pack_count;
    int vec_end;     // vectors of every level occupy LDS [0, vec_end)
};

// LDS layout: the padded vectors b, x, t, u of (h + 4) x (w + 4) of every
// level (element (y, x) at [(y + 2) * pitch + x + 2]), top grid first; then
// the [13][n] stencils of the levels that keep them in LDS and the dense
// inverse, in the order of the hierarchy's packed copy.
template <typename T>
struct TailView {
    int h, w, n, pitch, psize;
    const T* coef;
    T* b;
    T* x;
    T* t;
    T* u;
    __device__ __forceinline__ int at(int y, int x_) const { return (y + 2) * pitch + x_ + 2; }
};

__host__ __device__ inline int tail_psize(int h, int w) { return (h + 4) * (w + 4); }

template <typename T>
__device__ __forceinline__ TailView<T> tail_view(T* lds, const TailLevel<T>& L) {
    TailView<T> v;
    v.h = L.h;
    v.w = L.w;
    v.n = L.h * L.w;
    v.pitch = L.w + 4;
    v.psize = (L.h + 4) * v.pitch;
    v.b = lds + L.off;
    v.x = v.b + v.psize;
    v.t = v.x + v.psize;
    v.u = v.t + v.psize;
    v.coef = L.coef_lds ? static_cast<const T*>(lds + L.coef_off) : L.coef;
    return v;
}

// (A v) at (y, x): 25 taps, no bounds tests (zero halo).
template <typename T>
__device__ __forceinline__ T tail_Ax(const TailView<T>& V, const T* v, int y, int x) {
    const long long i = (long long)y * V.w + x;
    const T* c = v + y * V.pitch + x;  // top-left tap of the 5 x 5 neighbourhood
    T acc = 0;
#pragma unroll
    for (int s = 0; s < 25; ++s)
        acc += sym_coef(V.coef, (long long)V.n, V.w, i, s) * c[(s / 5) * V.pitch + s % 5];
    return acc;
}

// P e at fine point (y, x) from the padded coarse vector e (vcycle.py:6-29):
// weights 1/2 at coincident rows/columns, 1/4 at the two neighbours.
template <typename T>
__device__ __forceinline__ T tail_prolong(const TailView<T>& C, const T* e, int y, int x) {
    const int y0 = y >> 1, x0 = x >> 1;
    const bool oy = y & 1, ox = x & 1;
    const T* r0 = e + C.at(y0, x0);
    if (!oy && !ox) return T(0.25) * r0[0];
    if (!oy) return T(0.125) * (r0[0] + r0[1]);
    if (!ox) return T(0.125) * (r0[0] + r0[C.pitch]);
    return T(0.0625) * ((r0[0] + r0[1]) + (r0[C.pitch] + r0[C.pitch + 1]));
}

// basis >= 0: the right-hand side is the unit vector of the top grid's point
// `basis` instead of the top grid's b, and the correction is written to
// column `basis` of the row-major top-grid matrix basis_out (see
// build_tail_matrix).
template <typename T>
__device__ __forceinline__ void vc_tail_body(const TailParams<T>& P, int basis = -1,
                                             T* basis_out = nullptr) {
    extern __shared__ __align__(16) unsigned char tail_lds_raw[];
    T* lds = reinterpret_cast<T*>(tail_lds_raw);
    const int tid = threadIdx.x;
    const int nlev = P.nlev;
    // The top grid's stencil in registers when it is not staged in LDS (at
    // most two points per thread; float32 only, float64 would exceed the
    // registers of a 1024-thread workgroup), loaded first so it overlaps the
    // staging.
    const TailView<T> V0 = tail_view(lds, P.lv[0]);
    const bool top_reg = sizeof(T) == 4 && !P.lv[0].coef_lds && V0.n <= 2 * kTailBlock;
    T tc[2][25];
    if constexpr (sizeof(T) == 4) {
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            const int i = tid + k * kTailBlock;
            const long long ii = top_reg && i < V0.n ? i : 0;
#pragma unroll
            for (int s = 0; s < 25; ++s)
                tc[k][s] = top_reg ? sym_coef(V0.coef, (long long)V0.n, V0.w, ii, s) : T(0);
        }
    }
    // Stage the top right-hand side (with its zero halo) and the packed
    // stencils and inverse, and zero the other vectors (halos must read as
    // zero; interiors are written before they are read). A round issues all
    // of its global loads before its LDS stores, overlapping the register
    // stencil's loads and the zero stores; the stores go to disjoint LDS
    // ranges, so one barrier ends the phase.
    {
        constexpr int kStage = 16;
        const int rhs_n = V0.psize;
        const int total = rhs_n + P.pack_count;
        for (int base = 0; base < total; base += kStage * kTailBlock) {
            T v[kStage];
#pragma unroll
            for (int u = 0; u < kStage; ++u) {
                const int g = base + u * kTailBlock + tid;
                v[u] = T(0);
                if (g < rhs_n) {
                    const int y = g / V0.pitch - 2, x = g - (g / V0.pitch) * V0.pitch - 2;
                    if (y >= 0 && y < V0.h && x >= 0 && x < V0.w)
                        v[u] = basis < 0 ? P.lv[0].b[y * V0.w + x]
                                         : (y * V0.w + x == basis ? T(1) : T(0));
                } else if (g < total) {
                    v[u] = P.pack[g - rhs_n];
                }
            }
            if (base == 0)
                for (int i = rhs_n + tid; i < P.vec_end; i += kTailBlock) lds[i] = T(0);
#pragma unroll
            for (int u = 0; u < kStage; ++u) {
                const int g = base + u * kTailBlock + tid;
                if (g < rhs_n)
                    lds[g] = v[u];
                else if (g < total)
                    lds[P.pack_off + g - rhs_n] = v[u];
            }
        }
    }
    __syncthreads();
    // The pre-smoothed iterate of every level ends in .x or .t depending on
    // the number of sweeps (one buffer swap per sweep after the first).
    const bool pre_in_t = P.pre >= 2 && ((P.pre - 1) & 1);
    const T omega = P.omega;
    // A thread's points of level V; for the top grid with its stencil in
    // registers (reg) the point's register slot k is a compile-time index.
    auto for_points = [&](const TailView<T>& V, bool reg, auto&& body) {
        if (reg) {
#pragma unroll
            for (int k = 0; k < 2; ++k) {
                const int i = tid + k * kTailBlock;
                if (i < V.n) {
                    const int y = i / V.w;
                    body(k, i, y, i - y * V.w);
                }
            }
        } else {
            for (int i = tid; i < V.n; i += kTailBlock) {
                const int y = i / V.w;
                body(0, i, y, i - y * V.w);
            }
        }
    };
    auto Ax = [&](const TailView<T>& V, bool reg, int k, const T* v, int y, int x) -> T {
        if (sizeof(T) == 4 && reg) {
            const T* c = v + y * V.pitch + x;
            T acc = 0;
#pragma unroll
            for (int s = 0; s < 25; ++s) acc += tc[k][s] * c[(s / 5) * V.pitch + s % 5];
            return acc;
        }
        return tail_Ax(V, v, y, x);
    };
    auto diag = [&](const TailView<T>& V, bool reg, int k, int i) -> T {
        return (sizeof(T) == 4 && reg) ? tc[k][12] : V.coef[i];
    };
    // Descend: pre-smooth, residual, restrict.
    for (int l = 0; l + 1 < nlev; ++l) {
        const TailView<T> V = tail_view(lds, P.lv[l]);
        const bool reg = l == 0 && top_reg;
        T* cur = V.x;
        T* nxt = V.t;
        if (P.pre > 0) {
            for_points(V, reg, [&](int k, int i, int y, int x) {
                const int p = V.at(y, x);
                cur[p] = omega * V.b[p] / diag(V, reg, k, i);
            });
            for (int q = 1; q < P.pre; ++q) {
                __syncthreads();
                for_points(V, reg, [&](int k, int i, int y, int x) {
                    const int p = V.at(y, x);
                    nxt[p] = cur[p] + omega * (V.b[p] - Ax(V, reg, k, cur, y, x)) /
                                          diag(V, reg, k, i);
                });
                T* sw = cur;
                cur = nxt;
                nxt = sw;
            }
        }
        __syncthreads();
        for_points(V, reg, [&](int k, int i, int y, int x) {
            const int p = V.at(y, x);
            V.u[p] = V.b[p] - Ax(V, reg, k, cur, y, x);
        });
        __syncthreads();
        const TailView<T> C = tail_view(lds, P.lv[l + 1]);
        for (int I = tid; I < C.n; I += kTailBlock) {
            const int Y = I / C.w, X = I - Y * C.w;
            const T* r0 = V.u + V.at(2 * Y - 1, 2 * X - 1);
            const T* r1 = r0 + V.pitch;
            const T* r2 = r1 + V.pitch;
            const T s = (r0[0] + T(2) * r0[1] + r0[2]) + T(2) * (r1[0] + T(2) * r1[1] + r1[2]) +
                        (r2[0] + T(2) * r2[1] + r2[2]);
            C.b[C.at(Y, X)] = s * T(0.0625);
        }
        __syncthreads();
    }
    // Direct solve on the last grid: x = inv(A) b.
    {
        const TailView<T> C = tail_view(lds, P.lv[nlev - 1]);
        T* bc = C.u;  // contiguous copy of the right-hand side
        for (int j = tid; j < C.n; j += kTailBlock) {
            const int y = j / C.w;
            bc[j] = C.b[C.at(y, j - y * C.w)];
        }
        __syncthreads();
        const T* inv = P.inv_off >= 0 ? lds + P.inv_off : P.inv;
        for (int i = tid; i < C.n; i += kTailBlock) {
            T acc = 0;
            for (int j = 0; j < C.n; ++j) acc += inv[(long long)i * C.n + j] * bc[j];
            const int y = i / C.w;
            C.x[C.at(y, i - y * C.w)] = acc;
        }
        __syncthreads();
    }
    // Ascend: prolongate the correction and post-smooth.
    const T* e = tail_view(lds, P.lv[nlev - 1]).x;
    for (int l = nlev - 2; l >= 0; --l) {
        const TailView<T> V = tail_view(lds, P.lv[l]);
        const TailView<T> C = tail_view(lds, P.lv[l + 1]);
        const bool reg = l == 0 && top_reg;
        const T* xp = pre_in_t ? V.t : V.x;
        T* a = pre_in_t ? V.x : V.t;  // must not overwrite the pre-smoothed iterate
        for (int i = tid; i < V.n; i += kTailBlock) {
            const int y = i / V.w, x = i - y * V.w;
            const int p = V.at(y, x);
            a[p] = xp[p] + tail_prolong(C, e, y, x);
        }
        T* cur = a;
        T* nxt = V.u;
        for (int q = 0; q < P.post; ++q) {
            __syncthreads();
            for_points(V, reg, [&](int k, int i, int y, int x) {
                const int p = V.at(y, x);
                nxt[p] = cur[p] + omega * (V.b[p] - Ax(V, reg, k, cur, y, x)) / diag(V, reg, k, i);
            });
            T* sw = cur;
            cur = nxt;
            nxt = sw;
        }
        __syncthreads();
        e = cur;
    }
    // The top tail grid's correction goes back to global memory.
    {
        const TailView<T> V = tail_view(lds, P.lv[0]);
        for (int i = tid; i < V.n; i += kTailBlock) {
            const int y = i / V.w;
            if (basis < 0)
                P.lv[0].x[i] = e[V.at(y, i - y * V.w)];
            else
                basis_out[(size_t)i * V.n + basis] = e[V.at(y, i - y * V.w)];
        }
    }
}

// Work that rides in the tail's launch on the CUs the tail leaves idle: the
// CG solution update x += alpha p (p in the tail's type, x in float64 or in
// the tail's type, alpha of the current iteration from the solve state),
// streamed by workgroups 1.. of the grid while workgroup 0 runs the tail.
// The V-cycle never reads x, and the solve's freeze flag stops both alike.
struct TailSide {
    void* x = nullptr;
    const void* p = nullptr;
    long long n = 0;
    const CgState* st = nullptr;
    int x_f64 = 1;  // x is float64 (else the tail's type)
};
constexpr unsigned kTailSideBlocks = 255;

// x += alpha p over workgroup b of nb workgroups of BS threads (the
// arithmetic of k_cg_update_pq), four consecutive elements per access when
// both vectors are aligned for it.
template <typename TX, typename T, int BS>
__device__ __forceinline__ void tail_side_update_n(TX* __restrict__ x, const T* __restrict__ p,
                                                   long long n, double alpha, unsigned b,
                                                   unsigned nb) {
    const T alphaT = (T)alpha;
    const auto upd = [&](TX xv, T pv) -> TX {
        if constexpr (std::is_same<TX, T>::value)
            return xv + alphaT * pv;
        else
            return xv + alpha * (TX)pv;
    };
    constexpr int U = 2;
    const bool vec = ((reinterpret_cast<uintptr_t>(x) % alignof(Vec4<TX>)) |
                      (reinterpret_cast<uintptr_t>(p) % alignof(Vec4<T>))) == 0;
    const long long nq = vec ? n / 4 : 0;
    const long long stride = (long long)nb * BS;
    Vec4<TX>* __restrict__ x4 = reinterpret_cast<Vec4<TX>*>(x);
    const Vec4<T>* __restrict__ p4 = reinterpret_cast<const Vec4<T>*>(p);
    for (long long q0 = (long long)b * BS + threadIdx.x; q0 < nq; q0 += U * stride) {
        Vec4<TX> xv[U];
        Vec4<T> pv[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const long long q = q0 + u * stride;
            if (q < nq) {
                xv[u] = x4[q];
                pv[u] = p4[q];
            }
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const long long q = q0 + u * stride;
            if (q >= nq) continue;
            Vec4<TX> o;
            o.a = upd(xv[u].a, pv[u].a);
            o.b = upd(xv[u].b, pv[u].b);
            o.c = upd(xv[u].c, pv[u].c);
            o.d = upd(xv[u].d, pv[u].d);
            x4[q] = o;
        }
    }
    // The elements after the last whole group (all of them when not aligned).
    for (long long i = 4 * nq + (long long)b * BS + threadIdx.x; i < n; i += stride)
        x[i] = upd(x[i], p[i]);
}

template <typename TX, typename T>
__device__ __forceinline__ void tail_side_update(TX* __restrict__ x, const T* __restrict__ p,
                                                 long long n, double alpha, unsigned b,
                                                 unsigned nb) {
    tail_side_update_n<TX, T, kTailBlock>(x, p, n, alpha, b, nb);
}

template <typename T, bool kSide>
__global__ __launch_bounds__(kTailBlock) void k_vc_tail(TailParams<T> P, const CgState* guard,
                                                        TailSide side) {
    if (solve_done(guard)) return;
    if constexpr (kSide) {
        if (blockIdx.x > 0) {
            const double alpha = side.st->alpha;
            const T* p = static_cast<const T*>(side.p);
            if (side.x_f64)
                tail_side_update<double, T>(static_cast<double*>(side.x), p, side.n, alpha,
                                            blockIdx.x - 1, gridDim.x - 1);
            else
                tail_side_update<T, T>(static_cast<T*>(side.x), p, side.n, alpha,
                                       blockIdx.x - 1, gridDim.x - 1);
            return;
        }
    }
    vc_tail_body<T>(P);
}

// The tail as a matrix. The tail is a fixed linear map from its top grid's
// right-hand side to its top grid's correction; k_vc_tail_basis applies it to
// every unit vector of the top grid (one workgroup each, the tail's own code)
// and so writes the map as a dense row-major matrix, once per hierarchy.
// Each V-cycle then applies the matrix (k_tail_matvec, one wavefront per row)
// instead of running the tail's chain of dependent phases in one workgroup.
// The result is the tail's up to rounding (the sums run in another order).
template <typename T>
__global__ __launch_bounds__(kTailBlock) void k_vc_tail_basis(TailParams<T> P, T* out) {
    vc_tail_body<T>(P, (int)blockIdx.x, out);
}

constexpr int kMatvecBlock = 256;
constexpr int kMatvecRowsPerBlock = kMatvecBlock / kWave;

template <typename T>
__device__ __forceinline__ T wave_sum1(T v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
    return v;
}

// x = M b for the rows x cols row-major M (rows <= kTailMaxTop): workgroups
// below row_blocks take four rows each, one per wavefront, in chunks of
// 64 x kMatvecU columns; the others stream the CG solution update beside
// them (TailSide, as in k_vc_tail).
constexpr int kMatvecU = 8;

template <typename T, bool kSide>
__global__ __launch_bounds__(kMatvecBlock) void k_tail_matvec(const T* __restrict__ M,
                                                              const T* __restrict__ b,
                                                              T* __restrict__ x, int rows,
                                                              int cols, unsigned row_blocks,
                                                              const CgState* guard,
                                                              TailSide side) {
    if (solve_done(guard)) return;
    if constexpr (kSide) {
        if (blockIdx.x >= row_blocks) {
            const double alpha = side.st->alpha;
            const T* p = static_cast<const T*>(side.p);
            const unsigned nb = gridDim.x - row_blocks, bi = blockIdx.x - row_blocks;
            if (side.x_f64)
                tail_side_update_n<double, T, kMatvecBlock>(static_cast<double*>(side.x), p,
                                                            side.n, alpha, bi, nb);
            else
                tail_side_update_n<T, T, kMatvecBlock>(static_cast<T*>(side.x), p, side.n,
                                                       alpha, bi, nb);
            return;
        }
    }
    const int lane = threadIdx.x % kWave;
    const int r = (int)blockIdx.x * kMatvecRowsPerBlock + (int)threadIdx.x / kWave;
    if (r >= rows) return;
    const T* row = M + (size_t)r * cols;
    T acc = 0;
    for (int j0 = 0; j0 < cols; j0 += kWave * kMatvecU) {
        T mv[kMatvecU], bv[kMatvecU];
#pragma unroll
        for (int u = 0; u < kMatvecU; ++u) {
            const int j = j0 + lane + u * kWave;
            mv[u] = j < cols ? row[j] : T(0);
            bv[u] = j < cols ? b[j] : T(0);
        }
#pragma unroll
        for (int u = 0; u < kMatvecU; ++u) acc = fmadd(mv[u], bv[u], acc);
    }
    acc = wave_sum1(acc);
    if (lane == 0) x[r] = acc;
}

// LDS elements of a tail over levels [start, L) with the top level's stencil
// staged or not, including the staged inverse when it fits.
inline long long tail_elems(const hm_vc* vc, size_t start, bool top_coef, bool* inv_staged) {
    long long e = 0;
    const size_t L = vc->levels.size();
    for (size_t l = start; l < L; ++l) {
        const VcLevel& V = vc->levels[l];
        e += 4LL * tail_psize(V.h, V.w);
        if (l > start || top_coef) e += (long long)kSymSlots * V.n;
    }
    const long long nn = (long long)vc->levels[L - 1].n * vc->levels[L - 1].n;
    const size_t es = vc->prec == HM_F64 ? sizeof(double) : sizeof(float);
    *inv_staged = (e + nn) * (long long)es <= kTailLdsBytes;
    if (*inv_staged) e += nn;
    return e;
}

inline bool tail_fits(const hm_vc* vc, size_t start, bool* top_coef, bool* inv_staged) {
    const size_t es = vc->prec == HM_F64 ? sizeof(double) : sizeof(float);
    if (vc->levels.size() - start > (size_t)kTailMaxLevels) return false;
    if (vc->levels[start].n > kTailMaxTop) return false;
    bool inv = false;
    if (tail_elems(vc, start, true, &inv) * (long long)es <= kTailLdsBytes) {
        *top_coef = true;
        *inv_staged = inv;
        return true;
    }
    if (tail_elems(vc, start, false, &inv) * (long long)es <= kTailLdsBytes) {
        *top_coef = false;
        *inv_staged = inv;
        return true;
    }
    return false;
}

// First level of the tail: the largest suffix of grids (never level 0) that
// fits the LDS budget, or L - 1 if even the two coarsest grids do not.
inline size_t vc_tail_start(const hm_vc* vc) {
    const size_t L = vc->levels.size();
    if (L <= 1) return L;
    if (vc->coarse_sweeps > 0) return L - 1;  // no direct solve to end a tail with
    size_t t = L - 1;
    bool tc, is;
    while (t > 1 && tail_fits(vc, t - 1, &tc, &is)) --t;
    return t;
}

// Whether a V-cycle of this hierarchy runs an LDS tail (which can carry
// TailSide work).
inline bool vc_has_tail(const hm_vc* vc) {
    const size_t L = vc->levels.size();
    return L > 1 && vc->coarse_sweeps == 0 && vc_tail_start(vc) < L - 1;
}

// Tail parameters of levels [start, L) of a hierarchy; returns the LDS bytes.
template <typename T>
size_t make_tail_params(const hm_vc* vc, size_t start, TailParams<T>& P) {
    P = TailParams<T>{};
    const size_t L = vc->levels.size();
    P.nlev = (int)(L - start);
    P.pre = vc->pre;
    P.post = vc->post;
    P.omega = (T)vc->omega;
    P.inv = pick<T>(vc->dense_inv64, vc->dense_inv32);
    bool top_coef = true, inv_staged = false;
    if (!tail_fits(vc, start, &top_coef, &inv_staged)) {
        top_coef = true;
        inv_staged = false;
    }
    int off = 0;
    for (size_t l = start; l < L; ++l) {  // vectors
        const VcLevel& V = vc->levels[l];
        TailLevel<T>& t = P.lv[l - start];
        t.h = V.h;
        t.w = V.w;
        t.coef = pick<T>(V.A->c64, V.A->c32);
        t.b = static_cast<T*>(V.b);
        t.x = static_cast<T*>(V.x);
        t.off = off;
        t.coef_lds = (l > start || top_coef) ? 1 : 0;
        off += 4 * tail_psize(V.h, V.w);
    }
    P.vec_end = off;
    P.pack_off = off;
    for (size_t l = start; l < L; ++l) {  // staged stencils
        TailLevel<T>& t = P.lv[l - start];
        t.coef_off = t.coef_lds ? off : -1;
        if (t.coef_lds) off += kSymSlots * (int)vc->levels[l].n;
    }
    P.inv_off = inv_staged ? off : -1;
    if (inv_staged) off += (int)(vc->levels[L - 1].n * vc->levels[L - 1].n);
    P.pack_count = off - P.pack_off;
    P.pack = static_cast<const T*>(vc->tail_pack);
    return (size_t)off * sizeof(T);
}

// Builds the hierarchy's packed tail contents for the tail at `start`.
template <typename T>
int build_tail_pack(hm_vc* vc, size_t start, hipStream_t s) {
    TailParams<T> P;
    make_tail_params<T>(vc, start, P);
    dev_free(vc->tail_pack);
    vc->tail_pack = nullptr;
    vc->tail_start = start;
    if (P.pack_count == 0) return HM_OK;
    HM_TRY(dev_alloc(&vc->tail_pack, sizeof(T) * (size_t)P.pack_count));
    T* pack = static_cast<T*>(vc->tail_pack);
    for (int l = 0; l < P.nlev; ++l) {
        const TailLevel<T>& t = P.lv[l];
        if (!t.coef_lds) continue;
        HM_HIP(hipMemcpyAsync(pack + (t.coef_off - P.pack_off), t.coef,
                              sizeof(T) * kSymSlots * (size_t)t.h * t.w,
                              hipMemcpyDeviceToDevice, s));
    }
    if (P.inv_off >= 0) {
        const size_t nn = (size_t)vc->levels.back().n * vc->levels.back().n;
        HM_HIP(hipMemcpyAsync(pack + (P.inv_off - P.pack_off), P.inv, sizeof(T) * nn,
                              hipMemcpyDeviceToDevice, s));
    }
    return HM_OK;
}

// Folds the restriction of the grid f above the tail (one damped-Jacobi
// sweep from zero, residual, full weighting; vcycle.py:62-80) into the tail
// matrix: F = Tm G, with Tm the nt x nt tail matrix and G the nt x nf map
// b_f -> b_t, so that x_t = F b_f replaces the restriction's launch and the
// tail's. Column j of G is the restriction of e_j - A (omega / D_j) e_j,
// whose residual lives on the 5 x 5 points around j, so it reaches at most
// the 4 x 4 coarse points from ((jx - 2) >> 1, (jy - 2) >> 1) on.
// k_tail_fold_g: one thread per (column j, coarse slot c) writes Gc[j][c].
template <typename T>
__global__ __launch_bounds__(256) void k_tail_fold_g(int ht, int wt,
                                                     const T* __restrict__ coef,
                                                     const T* __restrict__ iD, int hf, int wf,
                                                     T omega, T* __restrict__ Gc) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const int nf = hf * wf;
    if (t >= 16LL * nf) return;
    const int j = (int)(t / 16), c = (int)(t % 16);
    const int jy = j / wf, jx = j - jy * wf;
    const int kx = ((jx - 2) >> 1) + c % 4, ky = ((jy - 2) >> 1) + c / 4;
    T v = T(0);
    if (kx >= 0 && kx < wt && ky >= 0 && ky < ht && 2 * kx + 1 >= jx - 2 && 2 * kx - 1 <= jx + 2 &&
        2 * ky + 1 >= jy - 2 && 2 * ky - 1 <= jy + 2) {
        const T z0 = vc_jac0(omega, T(1), iD[j]);
        T r[3][3];
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
            for (int q = 0; q < 3; ++q) {
                const int iy = 2 * ky - 1 + a, ix = 2 * kx - 1 + q;
                const int dy = iy - jy, dx = ix - jx;
                T rv = T(0);
                if (dy >= -2 && dy <= 2 && dx >= -2 && dx <= 2 && iy >= 0 && iy < hf && ix >= 0 &&
                    ix < wf) {
                    const int i = iy * wf + ix;
                    const int s = (2 - dy) * 5 + (2 - dx);  // slot of row i that couples to j
                    rv = (i == j ? T(1) : T(0)) - sym_coef32(coef, nf, wf, i, s) * z0;
                }
                r[a][q] = rv;
            }
        v = vc_restrict9(&r[0][1], &r[1][1], &r[2][1]);
    }
    Gc[t] = v;
}

// k_tail_fold_rows: one workgroup per row i of F, F[i][j] = sum_c
// Tm[i][K(j, c)] Gc[j][c], with Tm's row i staged in LDS.
template <typename T>
__global__ __launch_bounds__(256) void k_tail_fold_rows(const T* __restrict__ Tm, int ht, int wt,
                                                        const T* __restrict__ Gc, int hf, int wf,
                                                        T* __restrict__ F) {
    __shared__ T row[kTailMaxTop];
    const int i = blockIdx.x;
    const int nt = ht * wt, nf = hf * wf;
    for (int k = threadIdx.x; k < nt; k += blockDim.x) row[k] = Tm[(size_t)i * nt + k];
    __syncthreads();
    for (int j = threadIdx.x; j < nf; j += blockDim.x) {
        const int jy = j / wf, jx = j - jy * wf;
        const int kx0 = (jx - 2) >> 1, ky0 = (jy - 2) >> 1;
        T acc = 0;
#pragma unroll
        for (int c = 0; c < 16; ++c) {
            const int kx = kx0 + c % 4, ky = ky0 + c / 4;
            if (kx >= 0 && kx < wt && ky >= 0 && ky < ht)
                acc = fmadd(row[ky * wt + kx], Gc[(size_t)j * 16 + c], acc);
        }
        F[(size_t)i * nf + j] = acc;
    }
}

// The tail of the hierarchy at `start` as a dense matrix (k_vc_tail_basis),
// in the hierarchy's precision; needs the packed tail contents. With a cycle
// of one pre-smoothing sweep and a grid f = start - 1 >= 1 above the tail,
// f's restriction is folded in (k_tail_fold_*) and the matrix maps b_f to the
// tail's correction.
template <typename T>
int build_tail_matrix(hm_vc* vc, size_t start, hipStream_t s) {
    TailParams<T> P;
    const size_t lds = make_tail_params<T>(vc, start, P);
    const long long n = vc->levels[start].n;
    dev_free(vc->tail_mat);
    vc->tail_mat = nullptr;
    vc->tail_rows = vc->tail_cols = 0;
    vc->tail_src = 0;
    if (n > kTailMaxTop) {
        set_error("internal: tail grid of %lld points", n);
        return HM_ERR_ARG;
    }
    void* tm = nullptr;
    HM_TRY(dev_alloc(&tm, sizeof(T) * (size_t)(n * n)));
    HM_TRY(set_max_dynamic_lds(reinterpret_cast<const void*>(&k_vc_tail_basis<T>),
                               kTailLdsBytes));
    k_vc_tail_basis<T><<<(unsigned)n, kTailBlock, lds, s>>>(P, static_cast<T*>(tm));
    HM_LAUNCH_CHECK();
#ifndef HM_TAIL_FOLD
#define HM_TAIL_FOLD 1
#endif
    if (HM_TAIL_FOLD && vc->pre == 1 && start >= 2) {
        const VcLevel& Lf = vc->levels[start - 1];
        const VcLevel& Lt = vc->levels[start];
        void* F = nullptr;
        void* Gc = nullptr;
        int rc = dev_alloc(&F, sizeof(T) * (size_t)(Lt.n * Lf.n));
        if (rc == HM_OK) rc = dev_alloc(&Gc, sizeof(T) * 16 * (size_t)Lf.n);
        if (rc != HM_OK) {
            dev_free(tm);
            dev_free(F);
            return rc;
        }
        k_tail_fold_g<T><<<div_up(16 * Lf.n, 256), 256, 0, s>>>(
            Lt.h, Lt.w, pick<T>(Lf.A->c64, Lf.A->c32), pick<T>(Lf.A->idiag64, Lf.A->idiag32),
            Lf.h, Lf.w, (T)vc->omega, static_cast<T*>(Gc));
        k_tail_fold_rows<T><<<(unsigned)Lt.n, 256, 0, s>>>(
            static_cast<const T*>(tm), Lt.h, Lt.w, static_cast<const T*>(Gc), Lf.h, Lf.w,
            static_cast<T*>(F));
        const hipError_t e = hipGetLastError();
        // Back to the pool: a later allocation reuses them in stream order.
        dev_free(tm);
        dev_free(Gc);
        if (e != hipSuccess) {
            dev_free(F);
            set_error("tail fold failed: %s", hipGetErrorString(e));
            return HM_ERR_HIP;
        }
        vc->tail_mat = F;
        vc->tail_rows = (int)Lt.n;
        vc->tail_cols = (int)Lf.n;
        vc->tail_src = start - 1;
        return HM_OK;
    }
    vc->tail_mat = tm;
    vc->tail_rows = vc->tail_cols = (int)n;
    vc->tail_src = start;
    return HM_OK;
}

constexpr int kVcTailMatrix = 4;  // launch structure bit (hm_vc::pairs): tail as a matrix

// Whether a cycle starting its tail at `tail` applies the tail matrix, and
// the grid whose right-hand side it reads (the descent stops above it).
inline bool vc_tail_matrix(const hm_vc* vc, size_t tail, size_t* src) {
    const bool on = (vc->pairs & kVcTailMatrix) && vc->tail_mat && tail == vc->tail_start &&
                    vc->tail_rows == (int)vc->levels[tail].n;
    if (on && src) *src = vc->tail_src;
    return on;
}

template <typename T>
int launch_tail(hm_vc* vc, size_t start, const CgState* guard, hipStream_t s,
                const TailSide* side = nullptr) {
    size_t src = start;
    if (vc_tail_matrix(vc, start, &src)) {
        const VcLevel& V = vc->levels[start];
        const T* b = static_cast<const T*>(vc->levels[src].b);
        const int rows = vc->tail_rows, cols = vc->tail_cols;
        const unsigned rb = div_up(rows, kMatvecRowsPerBlock);
        ProfScope prof(kProfDense, rows, sizeof(T) == 4, s);
        if (side && side->x) {
            if (!side->p || !side->st || side->n < 0) {
                set_error("internal: incomplete V-cycle tail side work");
                return HM_ERR_ARG;
            }
            const unsigned sb = div_up(side->n, (long long)kMatvecBlock * 8);
            k_tail_matvec<T, true><<<rb + sb, kMatvecBlock, 0, s>>>(
                static_cast<const T*>(vc->tail_mat), b, static_cast<T*>(V.x), rows, cols, rb,
                guard, *side);
        } else {
            k_tail_matvec<T, false><<<rb, kMatvecBlock, 0, s>>>(
                static_cast<const T*>(vc->tail_mat), b, static_cast<T*>(V.x), rows, cols, rb,
                guard, TailSide{});
        }
        HM_LAUNCH_CHECK();
        return HM_OK;
    }
    TailParams<T> P;
    const size_t lds = make_tail_params<T>(vc, start, P);
    if (start != vc->tail_start || (P.pack_count > 0 && !P.pack)) {
        set_error("V-cycle tail at level %zu has no packed contents", start);
        return HM_ERR_ARG;
    }
    HM_TRY(set_max_dynamic_lds(reinterpret_cast<const void*>(&k_vc_tail<T, false>),
                               kTailLdsBytes));
    HM_TRY(set_max_dynamic_lds(reinterpret_cast<const void*>(&k_vc_tail<T, true>),
                               kTailLdsBytes));
    ProfScope prof(kProfDense, vc->levels[start].n, sizeof(T) == 4, s);
    if (side && side->x) {
        if (!side->p || !side->st || side->n < 0) {
            set_error("internal: incomplete V-cycle tail side work");
            return HM_ERR_ARG;
        }
        k_vc_tail<T, true><<<1 + kTailSideBlocks, kTailBlock, lds, s>>>(P, guard, *side);
    } else {
        k_vc_tail<T, false><<<1, kTailBlock, lds, s>>>(P, guard, TailSide{});
    }
    HM_LAUNCH_CHECK();
    return HM_OK;
}

}  // namespace hm
