// Batched PCG: many independent lambda systems (one per image, sizes may
// differ) solved together. Every step of an iteration is one launch over the
// tiles (or vector blocks) of all images that still iterate: workgroup ->
// (image, tile) through a tile table, per-image operands through an entry
// table, per-image CG scalars and freeze flags (the convergence mask) in an
// array of CgState, and per-image reductions finalised by one workgroup per
// image. The iteration's launch sequence does not depend on data, so it is
// recorded once into a HIP graph of `graph_iters` iterations and replayed
// until every image has converged.
//
// The per-image algorithm is cg.py:4-117 with the V-cycle of
// vcycle.py:46-100 (the same kernels as the single-image path); the mixed
// form is the float64 refinement around float32 PCG of pcg_mixed.
#include <cmath>
#include <cstring>
#include <functional>
#include <type_traits>
#include <vector>

#include "common.h"
#include "engine.h"
#include "internal.h"
#include "vcycle.h"

namespace hm {

struct TileRef {
    unsigned img, local;  // image, tile (or vector block) within the image
};

template <class Op, class In, class Out>
struct BatchEntry {
    Op op;
    In in;
    Out out;
    int h, w;
    unsigned tiles_x;
    const CgState* guard;
};

template <typename T, class Op, class In, class Out>
__global__ __launch_bounds__(kBlock) void tile_kernel_batch(
    const BatchEntry<Op, In, Out>* __restrict__ ents, const TileRef* __restrict__ tiles,
    unsigned ntiles) {
    const unsigned t = xcd_tile(blockIdx.x, ntiles);
    const TileRef tr = tiles[t];
    const BatchEntry<Op, In, Out> E = ents[tr.img];
    // A finished image's tiles return before loading anything (tile_body
    // would issue its loads first): in a batch, images finish tens of
    // iterations apart and the launches keep covering them.
    if (solve_done(E.guard)) return;
    const int by = (int)(tr.local / E.tiles_x);
    const int bx = (int)(tr.local - (unsigned)by * E.tiles_x);
    tile_body<T, Op, In, Out>(E.op, E.in, E.out, E.h, E.w, false, bx, by, t);
}

// ---- vector kernels over the images' vector blocks (kBlock*kVecPerThread) --

template <typename T>
struct UpdEntry {
    T* x;
    const T* p;
    T* r;
    const T* q;
    long long n;
    CgState* st;
};

// x += alpha p; r -= alpha q; partial ||r||^2 at the block's slot (cg.py:90-95).
template <typename T>
__global__ __launch_bounds__(kBlock) void k_cg_update_batch(const UpdEntry<T>* __restrict__ ents,
                                                            const TileRef* __restrict__ blocks,
                                                            Reducer red) {
    const TileRef br = blocks[blockIdx.x];
    const UpdEntry<T> E = ents[br.img];
    if (solve_done(E.st)) return;
    const T alpha = (T)E.st->alpha;
    double acc = 0.0;
    const long long base = (long long)br.local * kBlock * kVecPerThread + threadIdx.x;
#pragma unroll
    for (int k = 0; k < kVecPerThread; ++k) {
        const long long i = base + (long long)k * kBlock;
        if (i < E.n) {
            E.x[i] += alpha * E.p[i];
            const T ri = E.r[i] - alpha * E.q[i];
            E.r[i] = ri;
            acc += (double)ri * (double)ri;
        }
    }
    double v[1] = {acc};
    block_partials_at<kBlock, 1>(v, red, blockIdx.x);
}

// Reliable-update forms (float32 CG, float64 solution): x (float64) +=
// alpha p; r -= alpha q; partial ||r||^2 at the block's slot.
struct UpdEntryX64 {
    double* x;
    const float* p;
    float* r;
    const float* q;
    long long n;
    CgState* st;
};

__global__ __launch_bounds__(kBlock) void k_update_x64_batch(const UpdEntryX64* __restrict__ ents,
                                                             const TileRef* __restrict__ blocks,
                                                             Reducer red) {
    const TileRef br = blocks[blockIdx.x];
    const UpdEntryX64 E = ents[br.img];
    if (solve_done(E.st)) return;
    const double alpha = E.st->alpha;
    const float alpha32 = (float)alpha;
    double acc = 0.0;
    const long long base = (long long)br.local * kBlock * kVecPerThread + threadIdx.x;
#pragma unroll
    for (int k = 0; k < kVecPerThread; ++k) {
        const long long i = base + (long long)k * kBlock;
        if (i < E.n) {
            E.x[i] += alpha * (double)E.p[i];
            const float ri = E.r[i] - alpha32 * E.q[i];
            E.r[i] = ri;
            acc += (double)ri * (double)ri;
        }
    }
    double v[1] = {acc};
    block_partials_at<kBlock, 1>(v, red, blockIdx.x);
}

// r32 = (float)r64 with partial ||r64||^2 at the update's slot of that
// vector block (lo[img] + local: the launch may cover a subset of the
// images), so the following finalize tests the true residual.
struct RepEntry {
    const double* r64;
    float* r32;
    long long n;
    const CgState* gate;  // done: no replacement for this image
};

__global__ __launch_bounds__(kBlock) void k_replace_batch(const RepEntry* __restrict__ ents,
                                                          const TileRef* __restrict__ blocks,
                                                          const unsigned* __restrict__ lo,
                                                          Reducer red) {
    const TileRef br = blocks[blockIdx.x];
    const RepEntry E = ents[br.img];
    if (solve_done(E.gate)) return;
    double acc = 0.0;
    const long long base = (long long)br.local * kBlock * kVecPerThread + threadIdx.x;
#pragma unroll
    for (int k = 0; k < kVecPerThread; ++k) {
        const long long i = base + (long long)k * kBlock;
        if (i < E.n) {
            const double v = E.r64[i];
            E.r32[i] = (float)v;
            acc += v * v;
        }
    }
    double v[1] = {acc};
    block_partials_at<kBlock, 1>(v, red, lo[br.img] + br.local);
}

// gate[i].done = 0 exactly for the images whose replacement is due at
// iteration k of the unit (at[i], set by the host between units) and that
// are still iterating.
__global__ void k_gate_batch(CgState* __restrict__ gate, const CgState* __restrict__ st,
                             const int* __restrict__ at, int k, int count) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) gate[i].done = (at[i] == k && !solve_done(st + i)) ? 0 : 1;
}

// Records ||r||^2 of the images replaced in this iteration (the true one).
__global__ void k_note_rr_batch(CgState* __restrict__ st, const CgState* __restrict__ gate,
                                int count) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count && !solve_done(gate + i) && !solve_done(st + i)) st[i].aux = st[i].rr;
}

template <typename T>
struct DotEntry {
    const T* a;
    const T* b;
    long long n;
};

template <typename T>
__global__ __launch_bounds__(kBlock) void k_dot_batch(const DotEntry<T>* __restrict__ ents,
                                                      const TileRef* __restrict__ blocks,
                                                      Reducer red) {
    const TileRef br = blocks[blockIdx.x];
    const DotEntry<T> E = ents[br.img];
    double acc = 0.0;
    const long long base = (long long)br.local * kBlock * kVecPerThread + threadIdx.x;
#pragma unroll
    for (int k = 0; k < kVecPerThread; ++k) {
        const long long i = base + (long long)k * kBlock;
        if (i < E.n) acc += (double)E.a[i] * (double)E.b[i];
    }
    double v[1] = {acc};
    block_partials_at<kBlock, 1>(v, red, blockIdx.x);
}

// Elementwise steps: dst = src, dst = 0, dst = src * scale, dst += scale * src.
enum EwKind { kEwCopy = 0, kEwZero = 1, kEwScale = 2, kEwAxpy = 3 };

template <typename TD, typename TS>
struct EwEntry {
    TD* dst;
    const TS* src;
    const double* scale;  // one value per image (device)
    long long n;
    const CgState* guard;
};

template <typename TD, typename TS, int KIND>
__global__ __launch_bounds__(kBlock) void k_ew_batch(const EwEntry<TD, TS>* __restrict__ ents,
                                                     const TileRef* __restrict__ blocks) {
    const TileRef br = blocks[blockIdx.x];
    const EwEntry<TD, TS> E = ents[br.img];
    if (solve_done(E.guard)) return;
    double sc = 0.0;
    if constexpr (KIND == kEwScale || KIND == kEwAxpy) sc = *E.scale;
    const long long base = (long long)br.local * kBlock * kVecPerThread + threadIdx.x;
#pragma unroll
    for (int k = 0; k < kVecPerThread; ++k) {
        const long long i = base + (long long)k * kBlock;
        if (i >= E.n) continue;
        if constexpr (KIND == kEwCopy) E.dst[i] = (TD)E.src[i];
        if constexpr (KIND == kEwZero) E.dst[i] = TD(0);
        if constexpr (KIND == kEwScale) E.dst[i] = (TD)((double)E.src[i] * sc);
        if constexpr (KIND == kEwAxpy) E.dst[i] += (TD)(sc * (double)E.src[i]);
    }
}

// Per-image finalize: workgroup i sums image i's partials (slot 0 over
// lo[i]..lo[i+1]; for kDotUpdateRz also slot 2 over lo_upd[i]..lo_upd[i+1])
// and applies them to st[i].
template <int N>
__global__ __launch_bounds__(kFinBlock) void k_finalize_batch(Reducer red,
                                                              const unsigned* __restrict__ lo,
                                                              const unsigned* __restrict__ lo_upd,
                                                              CgState* st, int kind, int guarded) {
    const unsigned img = blockIdx.x;
    CgState* S = st + img;
    const bool done = guarded && solve_done(S);
    const double* src[N];
    unsigned cnt[N];
    const unsigned a = lo[img], b = lo[img + 1];
#pragma unroll
    for (int k = 0; k < N; ++k) {
        src[k] = red.partials + (size_t)k * red.max_blocks + a;
        cnt[k] = b - a;
    }
    if (kind == kDotUpdateRz) {
        const unsigned ua = lo_upd[img], ub = lo_upd[img + 1];
        src[0] = red.partials + 2 * (size_t)red.max_blocks + ua;
        cnt[0] = ub - ua;
        if (N > 1) {
            src[N - 1] = red.partials + a;
            cnt[N - 1] = b - a;
        }
    }
    if (done) return;  // uniform across the workgroup
    double acc[N];
    finalize_sums<N>(src, cnt, acc);
    if (threadIdx.x == 0) finalize_apply<N>(S, kind, acc);
}

// One workgroup per image: the LDS-resident tail of that image's V-cycle.
template <typename T>
__global__ __launch_bounds__(kTailBlock) void k_vc_tail_batch(const TailParams<T>* __restrict__ P,
                                                              const unsigned* __restrict__ img,
                                                              const CgState* st) {
    const unsigned i = img[blockIdx.x];
    if (solve_done(st + i)) return;
    vc_tail_body<T>(P[blockIdx.x]);
}

}  // namespace hm

// ===================================================================== batch

struct hm_batch {
    int count = 0;
    int prec = HM_F64;  // HM_F64: float64 PCG; HM_F32: mixed (float32 PCG in float64 refinement)
    int s0 = 0, sc = 1; // tile shapes: finest grid, Galerkin grids
    std::vector<hm_dia*> A;
    std::vector<hm_vc*> vc;
    std::vector<size_t> tail;  // first tail level per image
    hipStream_t stream = nullptr;  // the batch's own stream (graph capture)
    hipEvent_t ev_in = nullptr, ev_out = nullptr;
};

namespace hm {

using Step = std::function<int(hipStream_t)>;

// Device memory of one solve: uploaded tables, work vectors, states.
struct Arena {
    std::vector<void*> bufs;
    ~Arena() {
        for (void* p : bufs) dev_free(p);
    }
    template <typename X>
    int upload(const std::vector<X>& v, X** out) {
        *out = nullptr;
        if (v.empty()) return HM_OK;
        void* p = nullptr;
        HM_TRY(dev_alloc(&p, sizeof(X) * v.size()));
        bufs.push_back(p);
        HM_HIP(hipMemcpy(p, v.data(), sizeof(X) * v.size(), hipMemcpyHostToDevice));
        *out = static_cast<X*>(p);
        return HM_OK;
    }
    template <typename X>
    int alloc(size_t count, X** out) {
        void* p = nullptr;
        HM_TRY(dev_alloc(&p, sizeof(X) * (count ? count : 1)));
        bufs.push_back(p);
        *out = static_cast<X*>(p);
        return HM_OK;
    }
};

struct TileTable {
    std::vector<TileRef> refs;
    std::vector<unsigned> lo;  // per image start (count + 1 entries)
    TileRef* d_refs = nullptr;
    unsigned* d_lo = nullptr;
};

// Plain tiles (OX x OY) or restriction tiles (CX x CY coarse points) of the
// images in `which` at grid sizes (h[i], w[i]).
inline void build_tiles(TileTable& t, int count, const std::vector<char>& which,
                        const std::vector<int>& h, const std::vector<int>& w, int tx_div,
                        int ty_div, std::vector<unsigned>* tiles_x) {
    t.refs.clear();
    t.lo.assign(count + 1, 0);
    if (tiles_x) tiles_x->assign(count, 1);
    for (int i = 0; i < count; ++i) {
        t.lo[i] = (unsigned)t.refs.size();
        if (!which[i]) continue;
        const unsigned tx = div_up(w[i], tx_div), ty = div_up(h[i], ty_div);
        if (tiles_x) (*tiles_x)[i] = tx;
        for (unsigned k = 0; k < tx * ty; ++k) t.refs.push_back(TileRef{(unsigned)i, k});
    }
    t.lo[count] = (unsigned)t.refs.size();
}

template <typename T, class Op, class In, class Out>
int add_tile_step(std::vector<Step>& steps, Arena& ar,
                  const std::vector<BatchEntry<Op, In, Out>>& ents, const TileTable& tt) {
    if (tt.refs.empty()) return HM_OK;
    BatchEntry<Op, In, Out>* d = nullptr;
    HM_TRY(ar.upload(ents, &d));
    const TileRef* refs = tt.d_refs;
    const unsigned nt = (unsigned)tt.refs.size();
    steps.push_back([=](hipStream_t s) {
        tile_kernel_batch<T, Op, In, Out><<<nt, kBlock, 0, s>>>(d, refs, nt);
        HM_LAUNCH_CHECK();
        return (int)HM_OK;
    });
    return HM_OK;
}

template <int N>
void add_finalize_step(std::vector<Step>& steps, Reducer red, const unsigned* lo,
                       const unsigned* lo_upd, CgState* st, int count, int kind, int guarded) {
    steps.push_back([=](hipStream_t s) {
        k_finalize_batch<N><<<count, kFinBlock, 0, s>>>(red, lo, lo_upd, st, kind, guarded);
        HM_LAUNCH_CHECK();
        return (int)HM_OK;
    });
}

template <typename TD, typename TS, int KIND>
int add_ew_step(std::vector<Step>& steps, Arena& ar, const std::vector<EwEntry<TD, TS>>& ents,
                const TileTable& vt) {
    if (vt.refs.empty()) return HM_OK;
    EwEntry<TD, TS>* d = nullptr;
    HM_TRY(ar.upload(ents, &d));
    const TileRef* refs = vt.d_refs;
    const unsigned nb = (unsigned)vt.refs.size();
    steps.push_back([=](hipStream_t s) {
        k_ew_batch<TD, TS, KIND><<<nb, kBlock, 0, s>>>(d, refs);
        HM_LAUNCH_CHECK();
        return (int)HM_OK;
    });
    return HM_OK;
}

inline int run_steps(const std::vector<Step>& steps, hipStream_t s) {
    for (const Step& f : steps) HM_TRY(f(s));
    return HM_OK;
}

// Shapes and tables of one batch at one precision.
template <typename T, int S0, int SC>
struct BatchPlan {
    using FineOp = DiaOp<T, false>;
    using CoarseOp = std::conditional_t<SC == 1, DiaOp<T, true>, DiaOp<T, false>>;
    hm_batch* B;
    int count;
    std::vector<int> L;  // levels per image
    int maxtail = 0;
    // Tile tables: plain tiles per level (fine shape at level 0, coarse shape
    // after), restriction tiles per level, vector blocks of the finest grid.
    std::vector<TileTable> plain, restr;
    std::vector<std::vector<unsigned>> plain_tx, restr_tx;
    TileTable vec;
    Reducer red{};
    Arena* ar;

    int init(hm_batch* b, Arena& arena) {
        B = b;
        ar = &arena;
        count = b->count;
        L.resize(count);
        for (int i = 0; i < count; ++i) {
            L[i] = (int)b->vc[i]->levels.size();
            maxtail = std::max(maxtail, (int)b->tail[i]);
        }
        plain.resize(maxtail + 1);
        restr.resize(maxtail + 1);
        plain_tx.resize(maxtail + 1);
        restr_tx.resize(maxtail + 1);
        for (int l = 0; l <= maxtail; ++l) {
            std::vector<char> which(count, 0);
            std::vector<int> h(count, 1), w(count, 1), h2(count, 1), w2(count, 1);
            for (int i = 0; i < count; ++i) {
                // Level l runs tile kernels while l < tail (and l = 0 always, for the SpMV).
                if (l < (int)b->tail[i] || l == 0) {
                    which[i] = 1;
                    h[i] = b->vc[i]->levels[l].h;
                    w[i] = b->vc[i]->levels[l].w;
                    if (l + 1 < L[i]) {
                        h2[i] = b->vc[i]->levels[l + 1].h;
                        w2[i] = b->vc[i]->levels[l + 1].w;
                    }
                }
            }
            const int ox = l == 0 ? TileDims<S0>::X : TileDims<SC>::X;
            const int oy = l == 0 ? TileDims<S0>::Y : TileDims<SC>::Y;
            const int cx = l == 0 ? TileDims<S0>::CX : TileDims<SC>::CX;
            const int cy = l == 0 ? TileDims<S0>::CY : TileDims<SC>::CY;
            build_tiles(plain[l], count, which, h, w, ox, oy, &plain_tx[l]);
            std::vector<char> rw(count, 0);
            for (int i = 0; i < count; ++i) rw[i] = which[i] && l < (int)b->tail[i];
            build_tiles(restr[l], count, rw, h2, w2, cx, cy, &restr_tx[l]);
            HM_TRY(arena.upload(plain[l].refs, &plain[l].d_refs));
            HM_TRY(arena.upload(plain[l].lo, &plain[l].d_lo));
            HM_TRY(arena.upload(restr[l].refs, &restr[l].d_refs));
        }
        {
            std::vector<char> all(count, 1);
            std::vector<int> h(count), w(count);
            for (int i = 0; i < count; ++i) {
                h[i] = 1;
                w[i] = (int)std::min<long long>(b->A[i]->n, 1LL << 30);
            }
            build_tiles(vec, count, all, h, w, kBlock * kVecPerThread, 1, nullptr);
            HM_TRY(arena.upload(vec.refs, &vec.d_refs));
            HM_TRY(arena.upload(vec.lo, &vec.d_lo));
        }
        const size_t slots = std::max(plain[0].refs.size(), vec.refs.size()) + 64;
        double* parts = nullptr;
        HM_TRY(arena.alloc<double>(3 * slots, &parts));
        red.partials = parts;
        red.max_blocks = (int)slots;
        red.upd_blocks = 0;
        return HM_OK;
    }

    Reducer red_upd() const {
        Reducer r = red;
        r.partials += 2 * (size_t)red.max_blocks;
        return r;
    }

    FineOp fine_op(int i) const {
        const hm_dia* A = B->A[i];
        return FineOp{A->h, A->w, pick<T>(A->c64, A->c32)};
    }
    const T* inv_diag(int i, int l) const {
        const hm_dia* A = l == 0 ? B->A[i] : B->vc[i]->levels[l].A;
        return pick<T>(A->idiag64, A->idiag32);
    }
    CoarseOp coarse_op(int i, int l) const {
        const hm_dia* A = B->vc[i]->levels[l].A;
        return CoarseOp{A->h, A->w, pick<T>(A->c64, A->c32)};
    }

    // z = V-cycle(r) for every image with <r, z> finalised as fin_kind.
    int vcycle_steps(std::vector<Step>& steps, const std::vector<const T*>& r,
                     const std::vector<T*>& z, CgState* st, int fin_kind) {
        const int pre = B->vc[0]->pre, post = B->vc[0]->post;
        const T omega = (T)B->vc[0]->omega;
        const DotSink nodot{nullptr, Reducer{}, 0};
        std::vector<std::vector<T*>> pre_x(maxtail + 1, std::vector<T*>(count, nullptr));
        // Descend.
        for (int l = 0; l < maxtail; ++l) {
            if (l == 0)
                HM_TRY((descend_level<FineOp, S0>(steps, l, r, pre, omega, pre_x, st)));
            else
                HM_TRY((descend_level<CoarseOp, SC>(steps, l, r, pre, omega, pre_x, st)));
        }
        // Tails: one workgroup per image.
        {
            std::vector<TailParams<T>> P(count);
            std::vector<unsigned> img(count);
            size_t lds = 0;
            for (int i = 0; i < count; ++i) {
                lds = std::max(lds, make_tail_params<T>(B->vc[i], B->tail[i], P[i]));
                if (B->tail[i] != B->vc[i]->tail_start ||
                    (P[i].pack_count > 0 && !P[i].pack)) {
                    set_error("image %d: V-cycle tail without packed contents", i);
                    return HM_ERR_ARG;
                }
                img[i] = (unsigned)i;
            }
            TailParams<T>* dP = nullptr;
            unsigned* dimg = nullptr;
            HM_TRY(ar->upload(P, &dP));
            HM_TRY(ar->upload(img, &dimg));
            HM_TRY(set_max_dynamic_lds(reinterpret_cast<const void*>(&k_vc_tail_batch<T>),
                                       kTailLdsBytes));
            const int n = count;
            steps.push_back([=](hipStream_t s) {
                k_vc_tail_batch<T><<<n, kTailBlock, lds, s>>>(dP, dimg, st);
                HM_LAUNCH_CHECK();
                return (int)HM_OK;
            });
        }
        // Ascend.
        std::vector<T*> result(count, nullptr);
        for (int i = 0; i < count; ++i)
            result[i] = static_cast<T*>(B->vc[i]->levels[B->tail[i]].x);
        for (int l = maxtail - 1; l >= 0; --l) {
            if (l == 0)
                HM_TRY((ascend_level<FineOp, S0>(steps, l, r, z, post, pre, omega, pre_x,
                                                 result, st, fin_kind)));
            else
                HM_TRY((ascend_level<CoarseOp, SC>(steps, l, r, z, post, pre, omega, pre_x,
                                                   result, st, fin_kind)));
        }
        if (fin_kind == kDotUpdateRz)
            add_finalize_step<2>(steps, red, plain[0].d_lo, vec.d_lo, st, count, fin_kind, 1);
        else
            add_finalize_step<1>(steps, red, plain[0].d_lo, nullptr, st, count, fin_kind, 1);
        return HM_OK;
    }

    template <class Op>
    Op level_op(int i, int l) const {
        if constexpr (std::is_same<Op, FineOp>::value && std::is_same<Op, CoarseOp>::value) {
            return l == 0 ? fine_op(i) : coarse_op(i, l);
        } else if constexpr (std::is_same<Op, FineOp>::value) {
            return fine_op(i);
        } else {
            return coarse_op(i, l);
        }
    }

    const T* level_b(int i, int l, const std::vector<const T*>& r) const {
        return l == 0 ? r[i] : static_cast<const T*>(B->vc[i]->levels[l].b);
    }

    template <class Op, int S>
    int descend_level(std::vector<Step>& steps, int l, const std::vector<const T*>& r, int pre,
                      T omega, std::vector<std::vector<T*>>& pre_x, CgState* st) {
        const DotSink nodot{nullptr, Reducer{}, 0};
        using RestrictJ = BatchEntry<Op, InJac0<T>, OutRestrict<T, S>>;
        using RestrictV = BatchEntry<Op, InVec<T>, OutRestrict<T, S>>;
        using SmoothJ = BatchEntry<Op, InJac0<T>, OutSmooth<T, S>>;
        using SmoothV = BatchEntry<Op, InVec<T>, OutSmooth<T, S>>;
        std::vector<RestrictJ> rj;
        std::vector<RestrictV> rv;
        std::vector<SmoothJ> sj;
        std::vector<std::vector<SmoothV>> sv(pre > 2 ? pre - 2 : 0);
        std::vector<int> members;
        for (int i = 0; i < count; ++i)
            if (l < (int)B->tail[i]) members.push_back(i);
        for (int i = 0; i < count; ++i) {
            if (l >= (int)B->tail[i]) {
                // Placeholder entries keep the table indexed by image.
                rj.push_back(RestrictJ{});
                rv.push_back(RestrictV{});
                sj.push_back(SmoothJ{});
                for (auto& v : sv) v.push_back(SmoothV{});
                continue;
            }
            const VcLevel& lev = B->vc[i]->levels[l];
            const VcLevel& cl = B->vc[i]->levels[l + 1];
            const Op op = level_op<Op>(i, l);
            const T* b = level_b(i, l, r);
            const T* iD = inv_diag(i, l);
            const int h = lev.h, w = lev.w;
            const OutRestrict<T, S> rest{b, static_cast<T*>(cl.b), cl.w, cl.h, nullptr, nodot};
            const unsigned rtx = restr_tx[l][i], ptx = plain_tx[l][i];
            if (pre == 1) {
                rj.push_back(RestrictJ{op, InJac0<T>{b, iD, omega}, rest, h, w, rtx, st + i});
            } else {
                T* cur = static_cast<T*>(lev.x);
                T* nxt = static_cast<T*>(lev.t);
                sj.push_back(SmoothJ{op, InJac0<T>{b, iD, omega},
                                     OutSmooth<T, S>{b, iD, omega, cur, nodot}, h, w, ptx, st + i});
                for (int k = 2; k < pre; ++k) {
                    sv[k - 2].push_back(SmoothV{op, InVec<T>{cur},
                                                OutSmooth<T, S>{b, iD, omega, nxt, nodot}, h, w,
                                                ptx, st + i});
                    std::swap(cur, nxt);
                }
                rv.push_back(RestrictV{op, InVec<T>{cur}, rest, h, w, rtx, st + i});
                pre_x[l][i] = cur;
            }
        }
        if (members.empty()) return HM_OK;
        if (pre == 1) {
            HM_TRY((add_tile_step<T, Op, InJac0<T>, OutRestrict<T, S>>(steps, *ar, rj, restr[l])));
        } else {
            HM_TRY((add_tile_step<T, Op, InJac0<T>, OutSmooth<T, S>>(steps, *ar, sj, plain[l])));
            for (auto& v : sv)
                HM_TRY((add_tile_step<T, Op, InVec<T>, OutSmooth<T, S>>(steps, *ar, v, plain[l])));
            HM_TRY((add_tile_step<T, Op, InVec<T>, OutRestrict<T, S>>(steps, *ar, rv, restr[l])));
        }
        return HM_OK;
    }

    template <class Op, int S>
    int ascend_level(std::vector<Step>& steps, int l, const std::vector<const T*>& r,
                     const std::vector<T*>& z, int post, int pre, T omega,
                     const std::vector<std::vector<T*>>& pre_x, std::vector<T*>& result,
                     CgState* st, int fin_kind) {
        const DotSink nodot{nullptr, Reducer{}, 0};
        using BaseJ = InJac0<T>;
        using BaseV = InVec<T>;
        using FirstJ = BatchEntry<Op, InProlong<T, BaseJ>, OutSmooth<T, S>>;
        using FirstV = BatchEntry<Op, InProlong<T, BaseV>, OutSmooth<T, S>>;
        using Next = BatchEntry<Op, InVec<T>, OutSmooth<T, S>>;
        std::vector<FirstJ> fj;
        std::vector<FirstV> fv;
        std::vector<std::vector<Next>> nx(post > 1 ? post - 1 : 0);
        bool any = false;
        std::vector<T*> new_result = result;
        for (int i = 0; i < count; ++i) {
            if (l >= (int)B->tail[i]) {
                fj.push_back(FirstJ{});
                fv.push_back(FirstV{});
                for (auto& v : nx) v.push_back(Next{});
                continue;
            }
            any = true;
            const VcLevel& lev = B->vc[i]->levels[l];
            const VcLevel& cl = B->vc[i]->levels[l + 1];
            const Op op = level_op<Op>(i, l);
            const T* b = level_b(i, l, r);
            const T* iD = inv_diag(i, l);
            const bool top = l == 0;
            T* avail[2] = {nullptr, nullptr};
            int na = 0;
            for (void* cand : {lev.x, lev.t, lev.u})
                if (cand != nullptr && static_cast<T*>(cand) != pre_x[l][i] && na < 2)
                    avail[na++] = static_cast<T*>(cand);
            T* target = top ? z[i] : avail[0];
            T* other = top ? avail[0] : avail[1];
            if (target == nullptr || (post >= 2 && other == nullptr)) {
                set_error("internal: missing V-cycle work vectors");
                return HM_ERR_ARG;
            }
            const T* e = result[i];
            const unsigned ptx = plain_tx[l][i];
            const T* prev = nullptr;
            for (int k = 0; k < post; ++k) {
                T* outp = ((post - 1 - k) % 2 == 0) ? target : other;
                const DotSink ds = (top && k == post - 1) ? DotSink{st + i, red, fin_kind} : nodot;
                const OutSmooth<T, S> sm{b, iD, omega, outp, ds};
                if (k == 0) {
                    if (pre == 1) {
                        fj.push_back(FirstJ{op, InProlong<T, BaseJ>{BaseJ{b, iD, omega}, e, cl.w,
                                                                    cl.h},
                                            sm, lev.h, lev.w, ptx, st + i});
                        fv.push_back(FirstV{});
                    } else {
                        fv.push_back(FirstV{op, InProlong<T, BaseV>{BaseV{pre_x[l][i]}, e, cl.w,
                                                                    cl.h},
                                            sm, lev.h, lev.w, ptx, st + i});
                        fj.push_back(FirstJ{});
                    }
                } else {
                    nx[k - 1].push_back(Next{op, InVec<T>{prev}, sm, lev.h, lev.w, ptx, st + i});
                }
                prev = outp;
            }
            new_result[i] = target;
        }
        result = new_result;
        if (!any) return HM_OK;
        if (pre == 1)
            HM_TRY((add_tile_step<T, Op, InProlong<T, BaseJ>, OutSmooth<T, S>>(steps, *ar, fj,
                                                                               plain[l])));
        else
            HM_TRY((add_tile_step<T, Op, InProlong<T, BaseV>, OutSmooth<T, S>>(steps, *ar, fv,
                                                                               plain[l])));
        for (auto& v : nx)
            HM_TRY((add_tile_step<T, Op, InVec<T>, OutSmooth<T, S>>(steps, *ar, v, plain[l])));
        return HM_OK;
    }

    // One PCG iteration (p_cur -> p_next) for every image with the solution
    // in float64 (the reliable-update form, T = float): the SpMV, its
    // finalize and the update into `steps`, the preconditioner into `tail`
    // (residual replacements go between the two).
    int iteration_steps_x64(std::vector<Step>& steps, std::vector<Step>& tail,
                            const std::vector<double*>& x, const std::vector<T*>& rv,
                            const std::vector<T*>& zv, const std::vector<T*>& qv,
                            const std::vector<T*>& p_cur, const std::vector<T*>& p_next,
                            CgState* st) {
        static_assert(std::is_same<T, float>::value, "reliable updates run float32 CG");
        using ApplyE = BatchEntry<FineOp, InPupdate<T>, OutApply<T, S0>>;
        std::vector<ApplyE> ap;
        std::vector<UpdEntryX64> up;
        for (int i = 0; i < count; ++i) {
            const hm_dia* A = B->A[i];
            OutApply<T, S0> out{qv[i], DotSink{st + i, red, kDotPQ}, p_next[i]};
            ap.push_back(ApplyE{fine_op(i), InPupdate<T>{zv[i], p_cur[i], st + i}, out, A->h,
                                A->w, plain_tx[0][i], st + i});
            up.push_back(UpdEntryX64{x[i], p_next[i], rv[i], qv[i], A->n, st + i});
        }
        HM_TRY((add_tile_step<T, FineOp, InPupdate<T>, OutApply<T, S0>>(steps, *ar, ap,
                                                                         plain[0])));
        add_finalize_step<1>(steps, red, plain[0].d_lo, nullptr, st, count, kDotPQ, 1);
        {
            UpdEntryX64* d = nullptr;
            HM_TRY(ar->upload(up, &d));
            const TileRef* refs = vec.d_refs;
            const unsigned nb = (unsigned)vec.refs.size();
            const Reducer ru = red_upd();
            steps.push_back([=](hipStream_t s) {
                k_update_x64_batch<<<nb, kBlock, 0, s>>>(d, refs, ru);
                HM_LAUNCH_CHECK();
                return (int)HM_OK;
            });
        }
        std::vector<const T*> rc(rv.begin(), rv.end());
        return vcycle_steps(tail, rc, zv, st, kDotUpdateRz);
    }

    // One PCG iteration (p_cur -> p_next) for every image.
    int iteration_steps(std::vector<Step>& steps, const std::vector<T*>& x,
                        const std::vector<T*>& rv, const std::vector<T*>& zv,
                        const std::vector<T*>& qv, const std::vector<T*>& p_cur,
                        const std::vector<T*>& p_next, CgState* st) {
        using ApplyE = BatchEntry<FineOp, InPupdate<T>, OutApply<T, S0>>;
        std::vector<ApplyE> ap;
        std::vector<UpdEntry<T>> up;
        for (int i = 0; i < count; ++i) {
            const hm_dia* A = B->A[i];
            OutApply<T, S0> out{qv[i], DotSink{st + i, red, kDotPQ}, p_next[i]};
            ap.push_back(ApplyE{fine_op(i), InPupdate<T>{zv[i], p_cur[i], st + i}, out, A->h,
                                A->w, plain_tx[0][i], st + i});
            up.push_back(UpdEntry<T>{x[i], p_next[i], rv[i], qv[i], A->n, st + i});
        }
        HM_TRY((add_tile_step<T, FineOp, InPupdate<T>, OutApply<T, S0>>(steps, *ar, ap,
                                                                         plain[0])));
        add_finalize_step<1>(steps, red, plain[0].d_lo, nullptr, st, count, kDotPQ, 1);
        {
            UpdEntry<T>* d = nullptr;
            HM_TRY(ar->upload(up, &d));
            const TileRef* refs = vec.d_refs;
            const unsigned nb = (unsigned)vec.refs.size();
            const Reducer ru = red_upd();
            steps.push_back([=](hipStream_t s) {
                k_cg_update_batch<T><<<nb, kBlock, 0, s>>>(d, refs, ru);
                HM_LAUNCH_CHECK();
                return (int)HM_OK;
            });
        }
        std::vector<const T*> rc(rv.begin(), rv.end());
        return vcycle_steps(steps, rc, zv, st, kDotUpdateRz);
    }
};

// Pinned host mirror and device array of per-image solve states.
struct BatchStates {
    CgState* dev = nullptr;
    CgState* host = nullptr;
    int count = 0;
    BatchStates() = default;
    // Owns the pinned mirror: never copied (closures take dev/host pointers).
    BatchStates(const BatchStates&) = delete;
    BatchStates& operator=(const BatchStates&) = delete;
    int init(int n, Arena& ar) {
        count = n;
        HM_TRY(ar.alloc<CgState>(n, &dev));
        HM_HIP(hipHostMalloc(reinterpret_cast<void**>(&host), sizeof(CgState) * n,
                             hipHostMallocDefault));
        std::memset(host, 0, sizeof(CgState) * n);
        return HM_OK;
    }
    ~BatchStates() {
        if (host) (void)hipHostFree(host);
    }
    int push(hipStream_t s) {
        HM_HIP(hipMemcpyAsync(dev, host, sizeof(CgState) * count, hipMemcpyHostToDevice, s));
        return HM_OK;
    }
    int pull(hipStream_t s) {
        HM_HIP(hipMemcpyAsync(host, dev, sizeof(CgState) * count, hipMemcpyDeviceToHost, s));
        HM_HIP(hipStreamSynchronize(s));
        return HM_OK;
    }
};

// Work vectors and states of the batched PCG, reused by every inner solve.
template <typename T>
struct BatchWork {
    std::vector<T*> r, p, p2, q, z;
    BatchStates st;
    int init(hm_batch* B, Arena& ar) {
        const int count = B->count;
        r.resize(count);
        p.resize(count);
        p2.resize(count);
        q.resize(count);
        z.resize(count);
        for (int i = 0; i < count; ++i) {
            const size_t n = (size_t)B->A[i]->n;
            HM_TRY(ar.alloc<T>(n, &r[i]));
            HM_TRY(ar.alloc<T>(n, &p[i]));
            HM_TRY(ar.alloc<T>(n, &p2[i]));
            HM_TRY(ar.alloc<T>(n, &q[i]));
            HM_TRY(ar.alloc<T>(n, &z[i]));
        }
        return st.init(count, ar);
    }
};

// Batched PCG on the inner precision T: x_i = A_i^-1 b_i for the images with
// active[i], each to its own rtol * ||b_i||. stats[i] gets the counts.
template <typename T, int S0, int SC>
int batch_pcg(hm_batch* B, BatchPlan<T, S0, SC>& plan, BatchWork<T>& wk, Arena& ar,
              const std::vector<const T*>& b, const std::vector<T*>& x, double rtol, int maxiter,
              int graph_iters, const std::vector<char>& active, std::vector<hm_cg_stats>& stats,
              hipStream_t s) {
    const int count = B->count;
    const std::vector<T*>& r = wk.r;
    const std::vector<T*>& p = wk.p;
    const std::vector<T*>& p2 = wk.p2;
    const std::vector<T*>& q = wk.q;
    const std::vector<T*>& z = wk.z;
    BatchStates& st = wk.st;
    // ||b_i|| (cg.py:66).
    {
        std::vector<DotEntry<T>> de;
        for (int i = 0; i < count; ++i) de.push_back(DotEntry<T>{b[i], b[i], B->A[i]->n});
        DotEntry<T>* d = nullptr;
        HM_TRY(ar.upload(de, &d));
        k_dot_batch<T><<<(unsigned)plan.vec.refs.size(), kBlock, 0, s>>>(d, plan.vec.d_refs,
                                                                         plan.red);
        HM_LAUNCH_CHECK();
        k_finalize_batch<1><<<count, kFinBlock, 0, s>>>(plan.red, plan.vec.d_lo, nullptr, st.dev,
                                                        kDotAux, 0);
        HM_LAUNCH_CHECK();
        HM_TRY(st.pull(s));
    }
    for (int i = 0; i < count; ++i) {
        CgState& c = st.host[i];
        const double norm_b = std::sqrt(c.aux);
        std::memset(&c, 0, sizeof(CgState));
        c.atol = 0.0;
        c.rtol = rtol;
        c.norm_b = norm_b;
        c.rr = norm_b * norm_b;
        // x = 0, r = b: an inactive image or a zero right-hand side is done.
        c.done = (!active[i] || !(norm_b > 0.0)) ? 1 : 0;
        stats[i] = hm_cg_stats{0, c.done && active[i] ? 1 : 0, norm_b, norm_b};
    }
    HM_TRY(st.push(s));
    // x = 0, r = b, p = 0; z = M r with rz = <r, z> (cg.py:64-80).
    std::vector<Step> start;
    {
        std::vector<EwEntry<T, T>> zx, cr, zp;
        for (int i = 0; i < count; ++i) {
            const long long n = B->A[i]->n;
            zx.push_back(EwEntry<T, T>{x[i], nullptr, nullptr, n, nullptr});
            cr.push_back(EwEntry<T, T>{r[i], b[i], nullptr, n, nullptr});
            zp.push_back(EwEntry<T, T>{p[i], nullptr, nullptr, n, nullptr});
        }
        HM_TRY((add_ew_step<T, T, kEwZero>(start, ar, zx, plan.vec)));
        HM_TRY((add_ew_step<T, T, kEwCopy>(start, ar, cr, plan.vec)));
        HM_TRY((add_ew_step<T, T, kEwZero>(start, ar, zp, plan.vec)));
        std::vector<const T*> rc(r.begin(), r.end());
        HM_TRY(plan.vcycle_steps(start, rc, z, st.dev, kDotRzFirst));
    }
    HM_TRY(run_steps(start, s));
    // Iterations alternate the direction buffers, so a pair is the unit.
    std::vector<Step> pair;
    HM_TRY(plan.iteration_steps(pair, x, r, z, q, p, p2, st.dev));
    HM_TRY(plan.iteration_steps(pair, x, r, z, q, p2, p, st.dev));
    const int unit = graph_iters > 0 ? std::max(2, graph_iters & ~1) : 8;
    hipGraphExec_t exec = nullptr;
    hipGraph_t graph = nullptr;
    struct GraphGuard {
        hipGraph_t* g;
        hipGraphExec_t* e;
        ~GraphGuard() {
            if (*e) (void)hipGraphExecDestroy(*e);
            if (*g) (void)hipGraphDestroy(*g);
        }
    } gg{&graph, &exec};
    if (graph_iters > 0) {
        HM_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        int rc = HM_OK;
        for (int k = 0; k < unit / 2 && rc == HM_OK; ++k) rc = run_steps(pair, s);
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(s, &g);
        graph = g;
        HM_TRY(rc);
        HM_HIP(e);
        HM_HIP(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
    }
    int enqueued = 0;
    bool all_done = false;
    while (enqueued < maxiter && !all_done) {
        if (exec) {
            HM_HIP(hipGraphLaunch(exec, s));
        } else {
            for (int k = 0; k < unit / 2; ++k) HM_TRY(run_steps(pair, s));
        }
        enqueued += unit;
        HM_TRY(st.pull(s));
        all_done = true;
        for (int i = 0; i < count; ++i) {
            if (st.host[i].nonfinite) {
                set_error("non-finite value in the batched solve of image %d", i);
                return HM_ERR_VALUE;
            }
            if (!st.host[i].done) all_done = false;
        }
    }
    bool ok = true;
    for (int i = 0; i < count; ++i) {
        if (!active[i]) continue;
        const CgState& c = st.host[i];
        // Iterations past maxiter are frozen out only at unit granularity:
        // an image that needed more than maxiter counts as not converged.
        const bool conv = c.done && c.iters <= maxiter;
        stats[i].iterations = std::min(c.iters, maxiter);
        stats[i].norm_r = std::sqrt(c.rr);
        stats[i].converged = conv ? 1 : 0;
        ok = ok && conv;
    }
    if (!ok) {
        set_error("Conjugate gradient descent did not converge within %d iterations", maxiter);
        return HM_ERR_NOCONV;
    }
    return HM_OK;
}

// Mixed precision with reliable updates for a batch (see pcg_reliable): one
// float32 CG per image with the solution in float64. Units of kUnit
// iterations run back to back and the states are read after each. Each
// image's residual is replaced by its float64 residual (b - A x) at the
// iteration of the next unit where its norm is predicted, from the image's
// rate in the last unit, to have dropped by delta since the last
// replacement; those launches cover only the images due (tile and block
// lists compacted per unit and copied up once per unit), with a gate that
// also skips images that converged earlier in the unit. An image that
// converges on the recursion is confirmed on its float64 residual and, if
// that fails, restarted from it (beta = 0) through a V-cycle run on a
// separate state array, so the other images are not touched.
template <int S0, int SC>
int batch_reliable(hm_batch* B, Arena& ar, const double* const* b, double* const* x,
                   double rtol, double delta, int maxiter, int graph, hm_cg_stats* stats,
                   hipStream_t s) {
    constexpr int kUnit = 8;  // even: the direction buffers alternate
    const int count = B->count;
    BatchPlan<float, S0, SC> plan;
    HM_TRY(plan.init(B, ar));
    BatchPlan<double, S0, SC> plan64;  // tile tables of the float64 residual
    HM_TRY(plan64.init(B, ar));
    BatchWork<float> wk;
    HM_TRY(wk.init(B, ar));
    BatchStates& st = wk.st;
    BatchStates aux, rst;  // true residual norms; states of a restart V-cycle
    HM_TRY(aux.init(count, ar));
    HM_TRY(rst.init(count, ar));
    CgState* gate = nullptr;
    HM_TRY(ar.alloc<CgState>(count, &gate));
    int* at_dev = nullptr;  // replacement iteration per image in the running unit, or -1
    HM_TRY(ar.alloc<int>(count, &at_dev));
    std::vector<double*> r64(count);
    for (int i = 0; i < count; ++i) HM_TRY(ar.alloc<double>((size_t)B->A[i]->n, &r64[i]));
    std::vector<double*> xv(x, x + count);
    const DotSink nodot{nullptr, Reducer{}, 0};
    const TileRef* vrefs = plan.vec.d_refs;
    const unsigned* vlo = plan.vec.d_lo;
    const unsigned nvb = (unsigned)plan.vec.refs.size();
    const Reducer red = plan.red, red_upd = plan.red_upd();
    // Per-image entries of the float64 residual, the float32 copy and the norm.
    using ResE = BatchEntry<DiaOp<double, false>, InVec<double>, OutResidual<double, S0>>;
    ResE* d_res = nullptr;
    RepEntry* d_rep = nullptr;
    DotEntry<double>* d_dot = nullptr;
    {
        std::vector<ResE> re;
        std::vector<RepEntry> rp;
        std::vector<DotEntry<double>> de;
        for (int i = 0; i < count; ++i) {
            const hm_dia* A = B->A[i];
            re.push_back(ResE{DiaOp<double, false>{A->h, A->w, A->c64}, InVec<double>{x[i]},
                              OutResidual<double, S0>{b[i], r64[i], nodot}, A->h, A->w,
                              plan64.plain_tx[0][i], gate + i});
            rp.push_back(RepEntry{r64[i], wk.r[i], A->n, gate + i});
            de.push_back(DotEntry<double>{r64[i], r64[i], A->n});
        }
        HM_TRY(ar.upload(re, &d_res));
        HM_TRY(ar.upload(rp, &d_rep));
        HM_TRY(ar.upload(de, &d_dot));
    }
    // Compacted tile (float64 residual) and vector-block lists of image
    // subsets: host staging (pinned) and device copies, filled per unit.
    const TileTable& t64 = plan64.plain[0];
    const size_t cap_t = t64.refs.size(), cap_v = plan.vec.refs.size();
    TileRef *h_t = nullptr, *h_v = nullptr, *d_t = nullptr, *d_v = nullptr;
    HM_TRY(ar.alloc<TileRef>(cap_t, &d_t));
    HM_TRY(ar.alloc<TileRef>(cap_v, &d_v));
    HM_HIP(hipHostMalloc(reinterpret_cast<void**>(&h_t), sizeof(TileRef) * cap_t,
                         hipHostMallocDefault));
    struct PinGuard { void* p; ~PinGuard() { if (p) (void)hipHostFree(p); } } g_t{h_t};
    HM_HIP(hipHostMalloc(reinterpret_cast<void**>(&h_v), sizeof(TileRef) * cap_v,
                         hipHostMallocDefault));
    PinGuard g_v{h_v};
    int* at_host = nullptr;
    HM_HIP(hipHostMalloc(reinterpret_cast<void**>(&at_host), sizeof(int) * count,
                         hipHostMallocDefault));
    PinGuard g_a{at_host};
    struct Span { size_t t0, nt, v0, nv; };
    // Appends the lists of the images in `sel` and returns where they are.
    size_t fill_t = 0, fill_v = 0;
    auto append = [&](const std::vector<char>& sel) {
        Span sp{fill_t, 0, fill_v, 0};
        for (int i = 0; i < count; ++i) {
            if (!sel[i]) continue;
            for (unsigned k = t64.lo[i]; k < t64.lo[i + 1]; ++k) h_t[fill_t++] = t64.refs[k];
            for (unsigned k = plan.vec.lo[i]; k < plan.vec.lo[i + 1]; ++k)
                h_v[fill_v++] = plan.vec.refs[k];
        }
        sp.nt = fill_t - sp.t0;
        sp.nv = fill_v - sp.v0;
        return sp;
    };
    // Copies what was appended since `from` (one copy per list).
    auto upload = [&](size_t t_from, size_t v_from) -> int {
        if (fill_t > t_from)
            HM_HIP(hipMemcpyAsync(d_t + t_from, h_t + t_from, sizeof(TileRef) * (fill_t - t_from),
                                  hipMemcpyHostToDevice, s));
        if (fill_v > v_from)
            HM_HIP(hipMemcpyAsync(d_v + v_from, h_v + v_from, sizeof(TileRef) * (fill_v - v_from),
                                  hipMemcpyHostToDevice, s));
        return HM_OK;
    };
    // r64 = b - A x on the images of a span (gated), r32 = r64 on them.
    auto resid_step = [=](const Span sp) -> Step {
        const TileRef* refs = d_t + sp.t0;
        const unsigned nt = (unsigned)sp.nt;
        return [=](hipStream_t s2) {
            if (nt == 0) return (int)HM_OK;
            tile_kernel_batch<double, DiaOp<double, false>, InVec<double>,
                              OutResidual<double, S0>><<<nt, kBlock, 0, s2>>>(d_res, refs, nt);
            HM_LAUNCH_CHECK();
            return (int)HM_OK;
        };
    };
    auto copy_step = [=](const Span sp) -> Step {
        const TileRef* refs = d_v + sp.v0;
        const unsigned nv = (unsigned)sp.nv;
        return [=](hipStream_t s2) {
            if (nv == 0) return (int)HM_OK;
            k_replace_batch<<<nv, kBlock, 0, s2>>>(d_rep, refs, vlo, red_upd);
            HM_LAUNCH_CHECK();
            return (int)HM_OK;
        };
    };
    // The gate of iteration k (images due at k and still iterating) and the
    // note of the replaced ||r||^2 after the preconditioner.
    CgState* const st_dev = st.dev;
    auto gate_step = [=](int k) -> Step {
        const int n = count;
        const CgState* sd = st_dev;
        return [=](hipStream_t s2) {
            k_gate_batch<<<div_up(n, 256), 256, 0, s2>>>(gate, sd, at_dev, k, n);
            HM_LAUNCH_CHECK();
            return (int)HM_OK;
        };
    };
    auto note_step = [=]() -> Step {
        const int n = count;
        CgState* sd = st_dev;
        return [=](hipStream_t s2) {
            k_note_rr_batch<<<div_up(n, 256), 256, 0, s2>>>(sd, gate, n);
            HM_LAUNCH_CHECK();
            return (int)HM_OK;
        };
    };
    // Opens the gate of exactly the images in `sel` (rst is the staging).
    auto open_gate = [&](const std::vector<char>& sel) -> int {
        for (int i = 0; i < count; ++i) {
            rst.host[i] = CgState{};
            rst.host[i].done = sel[i] ? 0 : 1;
        }
        HM_TRY(rst.push(s));
        HM_HIP(hipMemcpyAsync(gate, rst.dev, sizeof(CgState) * count, hipMemcpyDeviceToDevice,
                              s));
        return HM_OK;
    };
    // ||b_i|| and the initial states.
    {
        std::vector<DotEntry<double>> de;
        for (int i = 0; i < count; ++i) de.push_back(DotEntry<double>{b[i], b[i], B->A[i]->n});
        DotEntry<double>* d = nullptr;
        HM_TRY(ar.upload(de, &d));
        k_dot_batch<double><<<nvb, kBlock, 0, s>>>(d, vrefs, red);
        HM_LAUNCH_CHECK();
        k_finalize_batch<1><<<count, kFinBlock, 0, s>>>(red, vlo, nullptr, st.dev, kDotAux, 0);
        HM_LAUNCH_CHECK();
        HM_TRY(st.pull(s));
    }
    std::vector<double> norm_b(count), rr_rep(count);
    for (int i = 0; i < count; ++i) {
        CgState& c = st.host[i];
        norm_b[i] = std::sqrt(c.aux);
        std::memset(&c, 0, sizeof(CgState));
        c.rtol = rtol;
        c.norm_b = norm_b[i];
        c.rr = norm_b[i] * norm_b[i];
        c.done = norm_b[i] > 0.0 ? 0 : 1;  // x = 0 solves b = 0
        rr_rep[i] = c.rr;
        stats[i] = hm_cg_stats{0, c.done, norm_b[i], norm_b[i]};
    }
    HM_TRY(st.push(s));
    // x = 0, p = 0, r32 = b; z = M r with rz = <r, z>.
    {
        std::vector<Step> start;
        std::vector<EwEntry<double, double>> zx;
        std::vector<EwEntry<float, float>> zp;
        std::vector<EwEntry<float, double>> cr;
        for (int i = 0; i < count; ++i) {
            const long long n = B->A[i]->n;
            zx.push_back(EwEntry<double, double>{x[i], nullptr, nullptr, n, nullptr});
            zp.push_back(EwEntry<float, float>{wk.p[i], nullptr, nullptr, n, nullptr});
            cr.push_back(EwEntry<float, double>{wk.r[i], b[i], nullptr, n, st.dev + i});
        }
        HM_TRY((add_ew_step<double, double, kEwZero>(start, ar, zx, plan.vec)));
        HM_TRY((add_ew_step<float, float, kEwZero>(start, ar, zp, plan.vec)));
        HM_TRY((add_ew_step<float, double, kEwCopy>(start, ar, cr, plan.vec)));
        std::vector<const float*> rc(wk.r.begin(), wk.r.end());
        HM_TRY(plan.vcycle_steps(start, rc, wk.z, st.dev, kDotRzFirst));
        HM_TRY(run_steps(start, s));
    }
    // Iteration k of a unit: head (SpMV, finalize, update) and preconditioner.
    std::vector<std::vector<Step>> head(kUnit), tail(kUnit);
    for (int k = 0; k < kUnit; ++k) {
        const bool even = k % 2 == 0;
        const std::vector<float*>& pc = even ? wk.p : wk.p2;
        const std::vector<float*>& pn = even ? wk.p2 : wk.p;
        HM_TRY(plan.iteration_steps_x64(head[k], tail[k], xv, wk.r, wk.z, wk.q, pc, pn,
                                        st.dev));
    }
    // graph != 0: each iteration's two fixed parts (head: SpMV, finalize,
    // update; tail: the V-cycle) are recorded once into a HIP graph each and
    // replayed; the replacements between them stay direct launches.
    std::vector<hipGraph_t> graphs;
    std::vector<hipGraphExec_t> execs;
    struct SegGuard {
        std::vector<hipGraph_t>* g;
        std::vector<hipGraphExec_t>* e;
        ~SegGuard() {
            for (hipGraphExec_t x : *e) (void)hipGraphExecDestroy(x);
            for (hipGraph_t x : *g) (void)hipGraphDestroy(x);
        }
    } seg_guard{&graphs, &execs};
    if (graph) {
        for (int k = 0; k < 2 * kUnit; ++k) {
            const std::vector<Step>& part = k % 2 == 0 ? head[k / 2] : tail[k / 2];
            HM_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
            const int rc = run_steps(part, s);
            hipGraph_t g = nullptr;
            const hipError_t e = hipStreamEndCapture(s, &g);
            if (g) graphs.push_back(g);
            HM_TRY(rc);
            HM_HIP(e);
            hipGraphExec_t x = nullptr;
            HM_HIP(hipGraphInstantiate(&x, g, nullptr, nullptr, 0));
            execs.push_back(x);
        }
    }
    auto segment = [&](int k, bool is_tail, std::vector<Step>& unit) {
        if (!graph) {
            const std::vector<Step>& part = is_tail ? tail[k] : head[k];
            unit.insert(unit.end(), part.begin(), part.end());
            return;
        }
        const hipGraphExec_t x = execs[2 * k + (is_tail ? 1 : 0)];
        unit.push_back([x](hipStream_t s2) {
            HM_HIP(hipGraphLaunch(x, s2));
            return (int)HM_OK;
        });
    };
    // Restart of a subset: z = M r32 with rz into rst (the others are done there).
    std::vector<Step> restart_vc;
    {
        std::vector<const float*> rc(wk.r.begin(), wk.r.end());
        HM_TRY(plan.vcycle_steps(restart_vc, rc, wk.z, rst.dev, kDotRzFirst));
    }
    std::vector<char> confirmed(count, 0);
    std::vector<int> rep_at(count, -1);  // none before a rate is known
    std::vector<int> iters_before(count, 0), it_prev(count, 0);
    std::vector<double> rr_prev(count), rho2(count, 0.0);
    for (int i = 0; i < count; ++i) {
        confirmed[i] = st.host[i].done;
        rr_prev[i] = rr_rep[i];
    }
    int enqueued = 0;
    while (enqueued < maxiter) {
        // This unit's steps, with the replacements of the images due.
        fill_t = fill_v = 0;
        std::vector<Step> unit;
        bool any_rep = false;
        for (int i = 0; i < count; ++i) {
            iters_before[i] = st.host[i].iters;
            at_host[i] = rep_at[i];
            any_rep = any_rep || rep_at[i] >= 0;
        }
        if (any_rep)
            HM_HIP(hipMemcpyAsync(at_dev, at_host, sizeof(int) * count, hipMemcpyHostToDevice,
                                  s));
        for (int k = 0; k < kUnit; ++k) {
            std::vector<char> due(count, 0);
            bool any = false;
            for (int i = 0; i < count; ++i) {
                due[i] = rep_at[i] == k;
                any = any || due[i];
            }
            segment(k, false, unit);
            if (any) {
                const Span sp = append(due);
                unit.push_back(gate_step(k));
                unit.push_back(resid_step(sp));
                unit.push_back(copy_step(sp));
            }
            segment(k, true, unit);
            if (any) unit.push_back(note_step());
        }
        HM_TRY(upload(0, 0));
        HM_TRY(run_steps(unit, s));
        enqueued += kUnit;
        HM_TRY(st.pull(s));
        std::vector<char> confirm(count, 0);
        bool any_confirm = false, all = true;
        for (int i = 0; i < count; ++i) {
            const CgState& c = st.host[i];
            if (c.nonfinite) {
                set_error("non-finite value in the batched solve of image %d", i);
                return HM_ERR_VALUE;
            }
            if (confirmed[i]) continue;
            const int rep_iter = rep_at[i] >= 0 ? iters_before[i] + rep_at[i] + 1 : -1;
            if (c.done) {
                if (c.iters == rep_iter) {
                    confirmed[i] = 1;  // tested on a true residual
                } else {
                    confirm[i] = 1;
                    any_confirm = true;
                }
            } else if (rep_iter >= 0) {
                rr_rep[i] = c.aux;  // noted in the replacement iteration
            }
        }
        if (any_confirm) {
            // True residuals of the images that converged on the recursion.
            fill_t = fill_v = 0;
            const Span sp = append(confirm);
            HM_TRY(upload(0, 0));
            HM_TRY(open_gate(confirm));
            HM_TRY(resid_step(sp)(s));
            k_dot_batch<double><<<(unsigned)sp.nv, kBlock, 0, s>>>(d_dot, d_v + sp.v0, red);
            HM_LAUNCH_CHECK();
            // The dot wrote partials at its compacted slots: give every image
            // its range within them.
            std::vector<unsigned> lo(count + 1, 0);
            {
                unsigned acc = 0;
                for (int i = 0; i < count; ++i) {
                    lo[i] = acc;
                    if (confirm[i]) acc += plan.vec.lo[i + 1] - plan.vec.lo[i];
                }
                lo[count] = acc;
            }
            unsigned* d_lo = nullptr;
            HM_TRY(ar.upload(lo, &d_lo));
            k_finalize_batch<1><<<count, kFinBlock, 0, s>>>(red, d_lo, nullptr, aux.dev,
                                                            kDotAux, 0);
            HM_LAUNCH_CHECK();
            HM_TRY(aux.pull(s));
            std::vector<char> restart(count, 0);
            bool any_restart = false;
            for (int i = 0; i < count; ++i) {
                if (!confirm[i]) continue;
                const double rr_true = aux.host[i].aux;
                if (std::sqrt(rr_true) < rtol * norm_b[i]) {
                    confirmed[i] = 1;
                    continue;
                }
                restart[i] = 1;
                any_restart = true;
                CgState& c = st.host[i];
                c.done = 0;
                c.rr = rr_true;
                c.beta = 0.0;
                rr_rep[i] = rr_prev[i] = rr_true;
                it_prev[i] = c.iters;
            }
            if (any_restart) {
                HM_TRY(st.push(s));
                // r32 = r64 on the restarted images, then z = M r32 for them.
                fill_t = fill_v = 0;
                const Span rs = append(restart);
                HM_TRY(upload(0, 0));
                HM_TRY(open_gate(restart));
                HM_TRY(copy_step(rs)(s));
                for (int i = 0; i < count; ++i) {
                    rst.host[i] = CgState{};
                    rst.host[i].done = restart[i] ? 0 : 1;
                }
                HM_TRY(rst.push(s));
                HM_TRY(run_steps(restart_vc, s));
                HM_TRY(rst.pull(s));
                for (int i = 0; i < count; ++i)
                    if (restart[i]) st.host[i].rz = rst.host[i].rz;
                HM_TRY(st.push(s));
            }
        }
        for (int i = 0; i < count; ++i) all = all && confirmed[i];
        if (all) break;
        // Replacement plan for the next unit: the iteration where the norm
        // is predicted to have dropped by delta since the last replacement.
        for (int i = 0; i < count; ++i) {
            const CgState& c = st.host[i];
            rep_at[i] = -1;
            if (confirmed[i] || c.done) continue;
            if (c.iters > it_prev[i] && c.rr < rr_prev[i])
                rho2[i] = std::pow(c.rr / rr_prev[i], 1.0 / (c.iters - it_prev[i]));
            rr_prev[i] = c.rr;
            it_prev[i] = c.iters;
            const double target = delta * delta * rr_rep[i];
            if (c.rr < target) {
                rep_at[i] = 0;
            } else if (rho2[i] > 0.0 && rho2[i] < 1.0) {
                const double j = std::ceil(std::log(target / c.rr) / std::log(rho2[i]));
                if (j <= kUnit) rep_at[i] = std::max(1, (int)j) - 1;
            }
        }
    }
    bool ok = true;
    for (int i = 0; i < count; ++i) {
        const CgState& c = st.host[i];
        stats[i].iterations = std::min(c.iters, maxiter);
        stats[i].norm_r = std::sqrt(c.rr);
        stats[i].converged = confirmed[i] && c.iters <= maxiter ? 1 : 0;
        ok = ok && stats[i].converged;
    }
    if (!ok) {
        set_error("Conjugate gradient descent did not converge within %d iterations", maxiter);
        return HM_ERR_NOCONV;
    }
    return HM_OK;
}

template <int S0, int SC>
int batch_solve_t(hm_batch* B, Arena& ar, const double* const* b, double* const* x, double rtol,
                  double inner_rtol, int max_outer, int maxiter, int graph_iters,
                  hm_cg_stats* stats, hipStream_t s) {
    const int count = B->count;
    std::vector<hm_cg_stats> st(count);
    std::vector<char> all(count, 1);
    if (B->prec == HM_F64) {
        BatchPlan<double, S0, SC> plan;
        HM_TRY(plan.init(B, ar));
        BatchWork<double> wk;
        HM_TRY(wk.init(B, ar));
        std::vector<const double*> bv(b, b + count);
        std::vector<double*> xv(x, x + count);
        const int rc = batch_pcg<double, S0, SC>(B, plan, wk, ar, bv, xv, rtol, maxiter,
                                                 graph_iters, all, st, s);
        for (int i = 0; i < count; ++i) stats[i] = st[i];
        return rc;
    }
    // Mixed: float64 refinement around float32 batched PCG (see pcg_mixed).
    BatchPlan<float, S0, SC> plan;
    HM_TRY(plan.init(B, ar));
    BatchWork<float> wk;
    HM_TRY(wk.init(B, ar));
    BatchPlan<double, S0, SC> plan64;  // tables of the float64 residual
    HM_TRY(plan64.init(B, ar));
    std::vector<double*> r64(count);
    std::vector<float*> r32(count), e32(count);
    for (int i = 0; i < count; ++i) {
        const size_t n = (size_t)B->A[i]->n;
        HM_TRY(ar.alloc<double>(n, &r64[i]));
        HM_TRY(ar.alloc<float>(n, &r32[i]));
        HM_TRY(ar.alloc<float>(n, &e32[i]));
    }
    double* scale_dev = nullptr;
    HM_TRY(ar.alloc<double>(2 * (size_t)count, &scale_dev));
    BatchStates ost;  // outer residual norms
    HM_TRY(ost.init(count, ar));
    // Steps: x = 0, r64 = b; ||r64||; r32 = r64 * s; x += s' e32; r64 = b - A x.
    std::vector<Step> init, norm, scale, axpy, resid;
    {
        std::vector<EwEntry<double, double>> zx, cb;
        std::vector<EwEntry<float, double>> sc;
        std::vector<EwEntry<double, float>> ax;
        std::vector<DotEntry<double>> de;
        using ResE = BatchEntry<DiaOp<double, false>, InVec<double>, OutResidual<double, S0>>;
        std::vector<ResE> re;
        const DotSink nodot{nullptr, Reducer{}, 0};
        for (int i = 0; i < count; ++i) {
            const hm_dia* A = B->A[i];
            const long long n = A->n;
            zx.push_back({x[i], nullptr, nullptr, n, nullptr});
            cb.push_back({r64[i], b[i], nullptr, n, nullptr});
            sc.push_back({r32[i], r64[i], scale_dev + i, n, nullptr});
            ax.push_back({x[i], e32[i], scale_dev + count + i, n, nullptr});
            de.push_back({r64[i], r64[i], n});
            re.push_back(ResE{DiaOp<double, false>{A->h, A->w, A->c64}, InVec<double>{x[i]},
                              OutResidual<double, S0>{b[i], r64[i], nodot}, A->h, A->w,
                              plan64.plain_tx[0][i], nullptr});
        }
        HM_TRY((add_ew_step<double, double, kEwZero>(init, ar, zx, plan.vec)));
        HM_TRY((add_ew_step<double, double, kEwCopy>(init, ar, cb, plan.vec)));
        HM_TRY((add_ew_step<float, double, kEwScale>(scale, ar, sc, plan.vec)));
        HM_TRY((add_ew_step<double, float, kEwAxpy>(axpy, ar, ax, plan.vec)));
        HM_TRY((add_tile_step<double, DiaOp<double, false>, InVec<double>,
                              OutResidual<double, S0>>(resid, ar, re, plan64.plain[0])));
        DotEntry<double>* d = nullptr;
        HM_TRY(ar.upload(de, &d));
        const TileRef* refs = plan.vec.d_refs;
        const unsigned nb = (unsigned)plan.vec.refs.size();
        const Reducer red = plan64.red;
        const unsigned* lo = plan.vec.d_lo;
        CgState* sd = ost.dev;
        norm.push_back([=](hipStream_t s2) {
            k_dot_batch<double><<<nb, kBlock, 0, s2>>>(d, refs, red);
            HM_LAUNCH_CHECK();
            k_finalize_batch<1><<<count, kFinBlock, 0, s2>>>(red, lo, nullptr, sd, kDotAux, 0);
            HM_LAUNCH_CHECK();
            return (int)HM_OK;
        });
    }
    HM_TRY(run_steps(init, s));
    std::vector<double> norm_b(count, 0.0), nr(count, 0.0);
    std::vector<char> conv(count, 0);
    std::vector<int> total(count, 0);
    std::vector<double> sc_host(2 * (size_t)count, 0.0);
    bool all_conv = false;
    for (int outer = 0; outer <= max_outer; ++outer) {
        HM_TRY(run_steps(norm, s));
        HM_TRY(ost.pull(s));
        all_conv = true;
        bool budget = false;
        for (int i = 0; i < count; ++i) {
            nr[i] = std::sqrt(ost.host[i].aux);
            if (outer == 0) norm_b[i] = nr[i];
            if (!conv[i] && (nr[i] < rtol * norm_b[i] || !(norm_b[i] > 0.0))) conv[i] = 1;
            if (!std::isfinite(nr[i])) {
                set_error("non-finite residual in the batched solve of image %d", i);
                return HM_ERR_VALUE;
            }
            all_conv = all_conv && conv[i];
            budget = budget || total[i] >= maxiter;
        }
        if (all_conv || outer == max_outer || budget) break;
        std::vector<char> active(count);
        for (int i = 0; i < count; ++i) {
            active[i] = !conv[i];
            sc_host[i] = conv[i] ? 0.0 : 1.0 / nr[i];
            sc_host[count + i] = conv[i] ? 0.0 : nr[i];
        }
        HM_HIP(hipMemcpy(scale_dev, sc_host.data(), sizeof(double) * 2 * count,
                         hipMemcpyHostToDevice));
        HM_TRY(run_steps(scale, s));
        std::vector<const float*> bv(r32.begin(), r32.end());
        std::vector<hm_cg_stats> inner(count);
        int maxin = 0;
        for (int i = 0; i < count; ++i) maxin = std::max(maxin, maxiter - total[i]);
        const int rc = batch_pcg<float, S0, SC>(B, plan, wk, ar, bv, e32, inner_rtol, maxin,
                                                graph_iters, active, inner, s);
        if (rc != HM_OK && rc != HM_ERR_NOCONV) return rc;
        for (int i = 0; i < count; ++i) total[i] += inner[i].iterations;
        HM_TRY(run_steps(axpy, s));
        HM_TRY(run_steps(resid, s));
    }
    for (int i = 0; i < count; ++i)
        stats[i] = hm_cg_stats{total[i], conv[i] ? 1 : 0, norm_b[i], nr[i]};
    if (!all_conv) {
        set_error("Conjugate gradient descent did not converge within %d iterations", maxiter);
        return HM_ERR_NOCONV;
    }
    return HM_OK;
}

}  // namespace hm

using namespace hm;

extern "C" {

int hm_batch_create(int count, hm_dia* const* ops, hm_vc* const* vcs, int precision,
                    int level0_shape, int coarse_shape, hm_batch** out) {
    if (!out || count < 1 || !ops || !vcs) {
        set_error("hm_batch_create: invalid argument");
        return HM_ERR_ARG;
    }
    if ((level0_shape != 0 && level0_shape != 1) || (coarse_shape != 0 && coarse_shape != 1)) {
        set_error("hm_batch_create: tile shapes are 0 (64x8) or 1 (32x8)");
        return HM_ERR_ARG;
    }
    const int want = precision == HM_F64 ? HM_F64 : HM_F32;
    int pre = -1, post = -1;
    double omega = 0.0;
    for (int i = 0; i < count; ++i) {
        hm_dia* A = ops[i];
        hm_vc* vc = vcs[i];
        if (!A || !vc) {
            set_error("hm_batch_create: null operator or V-cycle for image %d", i);
            return HM_ERR_ARG;
        }
        if (vc->prec != want) {
            set_error("hm_batch_create: image %d: the V-cycle precision must be %s", i,
                      want == HM_F64 ? "float64" : "float32 (mixed form)");
            return HM_ERR_ARG;
        }
        if (vc->cf != nullptr || vc->levels.size() < 2 || vc->levels[0].A != A ||
            vc->coarse_sweeps > 0) {
            set_error("hm_batch_create: image %d: the V-cycle must be built on this stored "
                      "operator with at least two grids and a direct coarsest solve", i);
            return HM_ERR_UNSUPPORTED;
        }
        if (vc_tail_start(vc) >= vc->levels.size() - 1) {
            set_error("hm_batch_create: image %d: the coarsest grids do not fit one workgroup",
                      i);
            return HM_ERR_UNSUPPORTED;
        }
        if (vc->pre < 1 || vc->post < 1) {
            set_error("hm_batch_create: batched V-cycles need at least one pre- and one "
                      "post-smoothing sweep");
            return HM_ERR_UNSUPPORTED;
        }
        if (i == 0) {
            pre = vc->pre;
            post = vc->post;
            omega = vc->omega;
        } else if (vc->pre != pre || vc->post != post || vc->omega != omega) {
            set_error("hm_batch_create: all images need the same V-cycle sweeps and omega");
            return HM_ERR_ARG;
        }
    }
    hm_batch* b = new hm_batch();
    b->count = count;
    b->prec = want;
    b->s0 = level0_shape;
    b->sc = coarse_shape;
    for (int i = 0; i < count; ++i) {
        b->A.push_back(ops[i]);
        b->vc.push_back(vcs[i]);
        b->tail.push_back(vc_tail_start(vcs[i]));
    }
    if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&b->ev_in, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&b->ev_out, hipEventDisableTiming) != hipSuccess) {
        set_error("hm_batch_create: stream or event creation failed");
        if (b->stream) (void)hipStreamDestroy(b->stream);
        if (b->ev_in) (void)hipEventDestroy(b->ev_in);
        delete b;
        return HM_ERR_HIP;
    }
    *out = b;
    return HM_OK;
}

int hm_batch_destroy(hm_batch* b) {
    if (!b) return HM_OK;
    if (b->stream) {
        (void)hipStreamSynchronize(b->stream);
        (void)hipStreamDestroy(b->stream);
    }
    if (b->ev_in) (void)hipEventDestroy(b->ev_in);
    if (b->ev_out) (void)hipEventDestroy(b->ev_out);
    delete b;
    return HM_OK;
}

int hm_batch_solve_reliable(hm_batch* b, const double* const* rhs, double* const* x,
                            double rtol, double delta, int maxiter, int graph,
                            hm_cg_stats* stats, void* stream) {
    if (!b || !rhs || !x || !stats || maxiter < 0 || !(delta > 0.0 && delta < 1.0) ||
        (graph != 0 && graph != 1)) {
        set_error("hm_batch_solve_reliable: invalid argument");
        return HM_ERR_ARG;
    }
    if (b->prec != HM_F32) {
        set_error("hm_batch_solve_reliable: the batch must be created for the mixed form");
        return HM_ERR_ARG;
    }
    for (int i = 0; i < b->count; ++i) {
        if (!rhs[i] || !x[i]) {
            set_error("hm_batch_solve_reliable: null vector for image %d", i);
            return HM_ERR_ARG;
        }
    }
    // The batch's own stream, ordered after and before the caller's stream.
    hipStream_t user = static_cast<hipStream_t>(stream);
    HM_HIP(hipEventRecord(b->ev_in, user));
    HM_HIP(hipStreamWaitEvent(b->stream, b->ev_in, 0));
    Arena ar;  // destroyed after the stream has drained
    int rc;
    if (b->s0 == 0 && b->sc == 0)
        rc = batch_reliable<0, 0>(b, ar, rhs, x, rtol, delta, maxiter, graph, stats,
                                       b->stream);
    else if (b->s0 == 0)
        rc = batch_reliable<0, 1>(b, ar, rhs, x, rtol, delta, maxiter, graph, stats,
                                       b->stream);
    else if (b->sc == 0)
        rc = batch_reliable<1, 0>(b, ar, rhs, x, rtol, delta, maxiter, graph, stats,
                                       b->stream);
    else
        rc = batch_reliable<1, 1>(b, ar, rhs, x, rtol, delta, maxiter, graph, stats,
                                       b->stream);
    (void)hipStreamSynchronize(b->stream);
    if (hipEventRecord(b->ev_out, b->stream) == hipSuccess)
        (void)hipStreamWaitEvent(user, b->ev_out, 0);
    return rc;
}

int hm_batch_solve(hm_batch* b, const double* const* rhs, double* const* x, double rtol,
                   double inner_rtol, int max_outer, int maxiter, int graph_iters,
                   hm_cg_stats* stats, void* stream) {
    if (!b || !rhs || !x || !stats || maxiter < 0 || graph_iters < 0) {
        set_error("hm_batch_solve: invalid argument");
        return HM_ERR_ARG;
    }
    for (int i = 0; i < b->count; ++i) {
        if (!rhs[i] || !x[i]) {
            set_error("hm_batch_solve: null vector for image %d", i);
            return HM_ERR_ARG;
        }
    }
    // The batch runs on its own stream (graphs cannot be captured on the
    // legacy default stream), ordered after and before the caller's stream.
    hipStream_t user = static_cast<hipStream_t>(stream);
    HM_HIP(hipEventRecord(b->ev_in, user));
    HM_HIP(hipStreamWaitEvent(b->stream, b->ev_in, 0));
    Arena ar;  // destroyed after the stream has drained (its buffers go back to the pool)
    int rc;
    if (b->s0 == 0 && b->sc == 0)
        rc = batch_solve_t<0, 0>(b, ar, rhs, x, rtol, inner_rtol, max_outer, maxiter, graph_iters,
                                 stats, b->stream);
    else if (b->s0 == 0)
        rc = batch_solve_t<0, 1>(b, ar, rhs, x, rtol, inner_rtol, max_outer, maxiter, graph_iters,
                                 stats, b->stream);
    else if (b->sc == 0)
        rc = batch_solve_t<1, 0>(b, ar, rhs, x, rtol, inner_rtol, max_outer, maxiter, graph_iters,
                                 stats, b->stream);
    else
        rc = batch_solve_t<1, 1>(b, ar, rhs, x, rtol, inner_rtol, max_outer, maxiter, graph_iters,
                                 stats, b->stream);
    (void)hipStreamSynchronize(b->stream);
    if (hipEventRecord(b->ev_out, b->stream) == hipSuccess)
        (void)hipStreamWaitEvent(user, b->ev_out, 0);
    return rc;
}

}  // extern "C"
