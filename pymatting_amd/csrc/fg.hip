// Multilevel foreground/background estimation on gfx950
// (pymatting/foreground/estimate_foreground_ml.py:61-183).
//
// The reference updates pixels in place in raster order (Gauss-Seidel): the
// new value of (y, x) uses the NEW values of (y, x-1) and (y-1, x) and the
// OLD values of (y, x+1) and (y+1, x). This file reproduces that order
// exactly with a strip-pipelined wavefront:
//   - a strip of 64 rows is swept by three 64-lane waves, one per colour
//     channel (each channel's F and B form their own 2x2 systems; only the
//     coefficients are shared). Lane y owns row y and at step t updates
//     column x = t - y, so lane y runs one column behind lane y-1. The new
//     value of the pixel above is lane y-1's result of the previous step,
//     and the old value of the pixel below is the "next old" register lane
//     y+1 loaded for its right neighbour; both move by cross-lane shuffles,
//     never through memory.
//   - strip s gets the new last row of strip s-1 through a halo buffer: each
//     value is stored with a 2-bit sweep tag in its two top bits (F and B are
//     clamped to [0, 1], whose float and double encodings leave those bits
//     zero), as a single relaxed device-scope atomic word, so a reader that
//     sees the current tag has the value without any fence or flag. Strip
//     s-1 reads the old first row of strip s before strip s may overwrite it:
//     strip s writes a column of its first row only after it has read that
//     column of strip s-1's last row, which strip s-1 writes after reading.
// Waits are bounded: a strip that waits too long raises a failure flag and
// finishes, so every wave exits. A level whose colours and coefficients fit
// in LDS (every level up to 32x32, where the reference makes its many small
// sweeps) runs instead in one workgroup with all its sweeps pipelined
// (k_fg_small).
//
// Arithmetic follows the types Numba gives the reference: with float32
// arrays, a0, a00, da and the b vector are float32 while a1 = 1.0 - a0 and
// everything derived from it are float64. Compiled with -ffp-contract=off.
#include <mutex>

#include "common.h"
#include "internal.h"

namespace hm {

namespace {

constexpr int kBlockMean = 256;
constexpr long long kSpinLimit = 1ll << 26;

template <typename S>
__device__ __forceinline__ S clamp01_store(double v) {
    v = v < 1.0 ? v : 1.0;  // max(0.0, min(1.0, F_c)), estimate_foreground_ml.py:171-172
    v = v > 0.0 ? v : 0.0;
    return (S)v;
}

// Coefficients of one pixel's 2x2 system (estimate_foreground_ml.py:132-165).
template <typename S>
struct PixelCoef {
    S da[4];       // left, right, up, down
    S b0[3], b1[3];
    double b00, b01, b11;
};

template <typename S>
__device__ __forceinline__ void pixel_coef(S a0, const S (&an)[4], const S* I, S reg, S gw,
                                           PixelCoef<S>& pc) {
    const double a1 = 1.0 - (double)a0;
    S a00 = a0 * a0;
    const double a01 = (double)a0 * a1;
    double a11 = a1 * a1;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        pc.b0[c] = a0 * I[c];
        pc.b1[c] = (S)(a1 * (double)I[c]);
    }
#pragma unroll
    for (int d = 0; d < 4; ++d) {
        const S grad = a0 > an[d] ? a0 - an[d] : an[d] - a0;  // abs(a0 - alpha[y2, x2])
        const S da = reg + gw * grad;
        pc.da[d] = da;
        a00 += da;
        a11 += (double)da;
    }
    const double det = (double)a00 * a11 - a01 * a01;
    const double inv_det = 1.0 / det;
    pc.b00 = inv_det * a11;
    pc.b01 = inv_det * -a01;
    pc.b11 = inv_det * (double)a00;
}

// The pixel's new F (res[0..2]) and B (res[3..5]) from its coefficients and
// the current values of its left, right, upper and lower neighbours
// (estimate_foreground_ml.py:158-175); shared by the sweep kernels so that
// they round alike.
// One colour channel: the channels' 2x2 systems share only the
// coefficients (v[0] is F, v[1] is B of the channel).
template <typename S>
__device__ __forceinline__ void fg_solve_channel(const S (&da)[4], S b0, S b1, double b00,
                                                 double b01, double b11, const S (&left)[2],
                                                 const S (&right)[2], const S (&up)[2],
                                                 const S (&down)[2], S (&res)[2]) {
    b0 += da[0] * left[0];
    b1 += da[0] * left[1];
    b0 += da[1] * right[0];
    b1 += da[1] * right[1];
    b0 += da[2] * up[0];
    b1 += da[2] * up[1];
    b0 += da[3] * down[0];
    b1 += da[3] * down[1];
    res[0] = clamp01_store<S>(b00 * (double)b0 + b01 * (double)b1);
    res[1] = clamp01_store<S>(b01 * (double)b0 + b11 * (double)b1);
}

template <typename S>
__device__ __forceinline__ void fg_solve_pixel(const S (&da)[4], const S* b0c, const S* b1c,
                                               double b00, double b01, double b11,
                                               const S (&left)[6], const S (&right)[6],
                                               const S (&up)[6], const S (&down)[6],
                                               S (&res)[6]) {
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const S l[2] = {left[c], left[3 + c]}, r[2] = {right[c], right[3 + c]};
        const S u[2] = {up[c], up[3 + c]}, d[2] = {down[c], down[3 + c]};
        S out[2];
        fg_solve_channel<S>(da, b0c[c], b1c[c], b00, b01, b11, l, r, u, d, out);
        res[c] = out[0];
        res[3 + c] = out[1];
    }
}

// Lane i gets lane i-1's value (shfl_up1) or lane i+1's (shfl_down1), the
// first / last lane keeps its own, as __shfl_up / __shfl_down by 1 over the
// wave; done with the wavefront-shift DPP moves (one VALU instruction per
// 32-bit word) instead of an LDS permute, since the sweep's step latency
// runs through them.
template <int Ctrl>
__device__ __forceinline__ float dpp_wave_move(float v) {
    const int b = __builtin_bit_cast(int, v);
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(b, b, Ctrl, 0xf, 0xf, false));
}
template <int Ctrl>
__device__ __forceinline__ double dpp_wave_move(double v) {
    const long long b = __builtin_bit_cast(long long, v);
    const int lo = (int)b, hi = (int)(b >> 32);
    const int rlo = __builtin_amdgcn_update_dpp(lo, lo, Ctrl, 0xf, 0xf, false);
    const int rhi = __builtin_amdgcn_update_dpp(hi, hi, Ctrl, 0xf, 0xf, false);
    return __builtin_bit_cast(double,
                              (long long)(((unsigned long long)(unsigned)rhi << 32) | (unsigned)rlo));
}
constexpr int kDppWaveShr1 = 0x138, kDppWaveShl1 = 0x130;
template <typename S>
__device__ __forceinline__ S shfl_up1(S v) { return dpp_wave_move<kDppWaveShr1>(v); }
template <typename S>
__device__ __forceinline__ S shfl_down1(S v) { return dpp_wave_move<kDppWaveShl1>(v); }

// Tagged halo words: the value's bit pattern with the sweep tag (1..3) in its
// two top bits.
template <typename S>
struct FgWord;
template <>
struct FgWord<float> {
    using U = unsigned int;
};
template <>
struct FgWord<double> {
    using U = unsigned long long;
};

template <typename S>
__device__ __forceinline__ typename FgWord<S>::U fg_tag(S v, unsigned tag) {
    using U = typename FgWord<S>::U;
    return __builtin_bit_cast(U, v) | ((U)tag << (8 * sizeof(U) - 2));
}
template <typename S>
__device__ __forceinline__ unsigned fg_word_tag(typename FgWord<S>::U b) {
    return (unsigned)(b >> (8 * sizeof(b) - 2));
}
template <typename S>
__device__ __forceinline__ S fg_untag(typename FgWord<S>::U b) {
    using U = typename FgWord<S>::U;
    return __builtin_bit_cast(S, (U)(b & ~((U)3 << (8 * sizeof(U) - 2))));
}

// Per-pixel coefficients of one level's 2x2 systems. They depend only on the
// level's alpha and colours, so they are computed once per level, in
// parallel, and every sweep of the level streams them. Planes: S [10][n]
// (da left/right/up/down, b0 rgb, b1 rgb), double [3][n] (b00, b01, b11).
template <typename S>
__global__ void k_fg_coef(const S* __restrict__ img, const S* __restrict__ alpha, int h, int w,
                          S reg, S gw, S* __restrict__ cs, double* __restrict__ cd) {
    const long long n = (long long)h * w;
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int y = (int)(i / w), x = (int)(i - (long long)y * w);
    const long long row = (long long)y * w;
    S an[4];
    an[0] = alpha[row + (x > 0 ? x - 1 : 0)];
    an[1] = alpha[row + (x + 1 < w ? x + 1 : x)];
    an[2] = alpha[(long long)(y > 0 ? y - 1 : 0) * w + x];
    an[3] = alpha[(long long)(y + 1 < h ? y + 1 : y) * w + x];
    const S I[3] = {img[i * 3], img[i * 3 + 1], img[i * 3 + 2]};
    PixelCoef<S> pc;
    pixel_coef<S>(alpha[i], an, I, reg, gw, pc);
#pragma unroll
    for (int d = 0; d < 4; ++d) cs[d * n + i] = pc.da[d];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        cs[(4 + c) * n + i] = pc.b0[c];
        cs[(7 + c) * n + i] = pc.b1[c];
    }
    cd[i] = pc.b00;
    cd[n + i] = pc.b01;
    cd[2 * n + i] = pc.b11;
}

// One in-place Gauss-Seidel sweep over an (h, w) level with the level's
// coefficient planes (k_fg_coef); F/B (h,w,3). halo[s][x][6] carries strip
// s's last row, tagged with this sweep's tag (1..3), to strip s+1.
//
// A workgroup of kFgThreads threads owns one strip. Waves 0-2 sweep it, one
// colour channel each (a channel's F and B; the channels' 2x2 systems share
// only their coefficients, so a step's work splits three ways with the same
// arithmetic). The other waves stage the inputs of the next chunk of C steps
// in LDS while the sweeping waves compute the current one (double
// buffering), and write the previous chunk's results from LDS to F/B with
// row-contiguous stores, so the sweeping waves only touch LDS (their lanes
// own different rows: their own stores would each touch 64 cache lines). Per
// lane and step the stage holds the pixel's coefficients and the old F/B one
// column to the right; the first loader wave also stages the new row above
// the strip (for lane 0, from strip s-1's tagged halo) and the old row below
// it (for the strip's last lane).
constexpr int kFgThreads = 1024;
constexpr int kFgSweepers = 3 * kWave;  // one wave per colour channel
constexpr int kFgLoaders = kFgThreads - kFgSweepers;
constexpr int kFgS = 16;  // S fields: da 4, b0 3, b1 3, old F/B to the right 6
constexpr int kFgD = 3;   // double fields: b00, b01, b11
constexpr int kFgPad = kWave + 1;  // lane stride (conflict-free staging stores)

template <typename S>
struct FgChunk {
    // Steps per chunk: as many as the double-buffered stage fits in LDS
    // (float32: 11 steps, 157 of the 160 KiB); each chunk ends with a
    // workgroup barrier, and longer chunks measured faster (4K float32
    // foreground 12.8 ms with 11 steps, 13.3 ms with 10, 17.0 ms with 5).
    static constexpr int C = sizeof(S) == 4 ? 11 : 5;
};

template <typename S>
struct FgStage {
    static constexpr int C = FgChunk<S>::C;
    S sv[2][C][kFgS][kFgPad];
    double dv[2][C][kFgD][kFgPad];
    S out[2][C][6][kFgPad];  // results, written back by the loader waves
    S up[2][C][6];
    S dn[2][C][6];
};

// Writes the results of chunk [t0, t0 + C) (buffer bf) to F/B; loader threads.
template <typename S>
__device__ __forceinline__ void fg_writeback(const FgStage<S>& st, int bf, int t0, int s, int lt,
                                             S* F, S* B, int h, int w) {
    constexpr int C = FgChunk<S>::C;
    for (int q = lt; q < kWave * C; q += kFgLoaders) {
        const int y = q / C, j = q - y * C;
        const int r = s * kWave + y, x = t0 + j - y;
        if (r >= h || x < 0 || x >= w) continue;
        const long long p = ((long long)r * w + x) * 3;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            F[p + c] = st.out[bf][j][c][y];
            B[p + c] = st.out[bf][j][3 + c][y];
        }
    }
}

// Stages the rows above and below the strip for chunk [t0, t0 + C) into
// buffer `bf`; run by the loader threads (lt = threadIdx.x - kFgSweepers).
// Returns false if the wait for strip s-1 timed out.
template <typename S>
__device__ __forceinline__ bool fg_stage_rows(FgStage<S>& st, int bf, int t0, int s, int lt,
                                              const S* F, const S* B, int h, int w,
                                              int last_lane, const typename FgWord<S>::U* halo,
                                              unsigned tag, long long& waited) {
    using W = typename FgWord<S>::U;
    constexpr int C = FgChunk<S>::C;
    bool ok = true;
    // Row above the strip (new values, from strip s-1's halo words, waiting
    // for the current tag) and row below it (old values). Wave 1 only.
    if (lt < kWave) {
        const int r0 = s * kWave;
        const bool is_up = lt < C, is_dn = lt >= 32 && lt < 32 + C;
        const int j = is_up ? lt : lt - 32;
        if (is_up && r0 > 0) {
            const int x = t0 + j;
            const int xc = x < 0 ? 0 : (x >= w ? w - 1 : x);
            const W* src = halo + ((long long)(s - 1) * w + xc) * 6;
            W v[6];
            while (true) {
                bool all = true;
#pragma unroll
                for (int q = 0; q < 6; ++q) {
                    v[q] = __hip_atomic_load(src + q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    all = all && fg_word_tag<S>(v[q]) == tag;
                }
                if (all) break;
                if (++waited > kSpinLimit) {
                    ok = false;
                    break;
                }
                __builtin_amdgcn_s_sleep(1);
            }
#pragma unroll
            for (int q = 0; q < 6; ++q) st.up[bf][j][q] = fg_untag<S>(v[q]);
        }
        if (is_dn) {
            const int x = t0 + j - last_lane;
            const int xc = x < 0 ? 0 : (x >= w ? w - 1 : x);
            const int rr = s * kWave + last_lane + 1;
            if (rr < h) {
                const long long q = ((long long)rr * w + xc) * 3;
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    st.dn[bf][j][c] = F[q + c];
                    st.dn[bf][j][3 + c] = B[q + c];
                }
            }
        }
    }
    return ok;
}

// A loader thread's per-lane inputs of one chunk (U (lane, step) pairs),
// held in registers between their loads and their LDS stores.
template <typename S>
struct FgLane {
    static constexpr int C = FgChunk<S>::C;
    static constexpr int U = (kWave * C + kFgLoaders - 1) / kFgLoaders;
    S sval[U][kFgS];
    double dval[U][kFgD];
};

// Loads chunk [t0, t0 + C)'s per-lane inputs (coefficients and the old F/B
// one column to the right) into registers; issued a chunk before they are
// stored, so their latency overlaps a whole chunk of sweeping.
template <typename S>
__device__ __forceinline__ void fg_load_lanes(FgLane<S>& L, int t0, int s, int lt,
                                              const S* __restrict__ cs,
                                              const double* __restrict__ cd, const S* F,
                                              const S* B, int h, int w) {
    constexpr int C = FgChunk<S>::C;
    constexpr int U = FgLane<S>::U;
    const long long n = (long long)h * w;
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const int q = lt + u * kFgLoaders;
        const bool used = q < kWave * C;
        const int y = used ? q / C : 0, j = used ? q - y * C : 0;
        const int r = s * kWave + y;
        const int rc = r < h ? r : h - 1;
        const long long row = (long long)rc * w;
        const int x = t0 + j - y;
        const int xc = x < 0 ? 0 : (x >= w ? w - 1 : x);
        const int xn = xc + 1 < w ? xc + 1 : xc;
        const long long i = row + xc;
        const long long pn = (row + (x + 1 >= 0 && x + 1 < w ? x + 1 : xn)) * 3;
#pragma unroll
        for (int f = 0; f < 10; ++f) L.sval[u][f] = cs[f * n + i];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            L.sval[u][10 + c] = F[pn + c];
            L.sval[u][13 + c] = B[pn + c];
        }
#pragma unroll
        for (int f = 0; f < kFgD; ++f) L.dval[u][f] = cd[f * n + i];
    }
}

// Stores loaded per-lane inputs into buffer bf.
template <typename S>
__device__ __forceinline__ void fg_store_lanes(FgStage<S>& st, int bf, int lt,
                                               const FgLane<S>& L) {
    constexpr int C = FgChunk<S>::C;
#pragma unroll
    for (int u = 0; u < FgLane<S>::U; ++u) {
        const int q = lt + u * kFgLoaders;
        if (q >= kWave * C) continue;
        const int y = q / C, j = q - y * C;
#pragma unroll
        for (int f = 0; f < kFgS; ++f) st.sv[bf][j][f][y] = L.sval[u][f];
#pragma unroll
        for (int f = 0; f < kFgD; ++f) st.dv[bf][j][f][y] = L.dval[u][f];
    }
}

template <typename S>
__global__ __launch_bounds__(kFgThreads) void k_fg_sweep(const S* __restrict__ cs,
                                                         const double* __restrict__ cd, S* F,
                                                         S* B, int h, int w,
                                                         typename FgWord<S>::U* halo,
                                                         unsigned tag, int* fail) {
    constexpr int C = FgChunk<S>::C;
    __shared__ FgStage<S> st;
    const int s = blockIdx.x;
    const bool computer = threadIdx.x < kFgSweepers;
    const int ch = threadIdx.x / kWave;         // sweeping waves: colour channel
    const int lane = threadIdx.x % kWave;       // sweeping waves: row in the strip
    const int lt = threadIdx.x - kFgSweepers;   // loaders
    const int last_lane = min(kWave, h - s * kWave) - 1;  // lane owning the strip's last row
    const int steps = w + kWave - 1;
    const int nchunks = (steps + C - 1) / C;
    long long waited = 0;
    bool ok = true;

    const int r = s * kWave + lane;
    const bool row_ok = r < h;
    const int rc = row_ok ? r : h - 1;
    const long long row = (long long)rc * w;
    // The channel's F and B (index 0 and 1) of the current pixel before
    // this sweep and of the previous pixel after it.
    S cur_old[2], prev_res[2];
    if (computer) {
        // Old values at column 0 (the first step's "current" pixel).
        cur_old[0] = F[row * 3 + ch];
        cur_old[1] = B[row * 3 + ch];
        prev_res[0] = cur_old[0];
        prev_res[1] = cur_old[1];
    }
    // Loaders: a chunk's per-lane inputs are loaded into registers one chunk
    // before they are stored into LDS.
    FgLane<S> lanes;
    if (!computer) {
        fg_load_lanes<S>(lanes, 0, s, lt, cs, cd, F, B, h, w);
        ok = fg_stage_rows<S>(st, 0, 0, s, lt, F, B, h, w, last_lane, halo, tag, waited);
        fg_store_lanes<S>(st, 0, lt, lanes);
        if (nchunks > 1) fg_load_lanes<S>(lanes, C, s, lt, cs, cd, F, B, h, w);
    }
    __syncthreads();
    for (int k = 0; k < nchunks; ++k) {
        const int bf = k & 1;
        const int t0 = k * C;
        if (computer) {
            const int t1 = min(t0 + C, steps);
            // The chunk's steps unrolled, the last chunk's surplus steps
            // predicated off (no branch), so each step's LDS reads, which do
            // not depend on earlier steps, can be issued ahead of the step
            // chain.
#pragma unroll
            for (int j = 0; j < C; ++j) {
                const int t = t0 + j;
                const bool step_ok = t < t1;  // uniform across the wave
                const int x = t - lane;
                const bool active = step_ok && row_ok && x >= 0 && x < w;
                const int xc = x < 0 ? 0 : (x >= w ? w - 1 : x);
                const S* sv = &st.sv[bf][j][0][lane];
                // This channel's fields: F at q = ch, B at q = 3 + ch.
                const int qf = ch, qb = 3 + ch;
                S next_old[2] = {sv[(10 + qf) * kFgPad], sv[(10 + qb) * kFgPad]};
                // The staged rows above and below the strip (read by every
                // lane, used by the first and last).
                const S up_s[2] = {st.up[bf][j][qf], st.up[bf][j][qb]};
                const S dn_s[2] = {st.dn[bf][j][qf], st.dn[bf][j][qb]};
                S left[2], right[2], up[2], down[2];
                S from_above[2], from_below[2];
#pragma unroll
                for (int v = 0; v < 2; ++v) {
                    from_above[v] = shfl_up1(prev_res[v]);    // lane-1's result at this column
                    from_below[v] = shfl_down1(next_old[v]);  // lane+1's old value at this column
                }
#pragma unroll
                for (int v = 0; v < 2; ++v) {
                    left[v] = x > 0 ? prev_res[v] : cur_old[v];
                    right[v] = x + 1 < w ? next_old[v] : cur_old[v];
                    up[v] = r == 0 ? cur_old[v] : (lane > 0 ? from_above[v] : up_s[v]);
                    down[v] = r >= h - 1 ? cur_old[v] : (lane < last_lane ? from_below[v] : dn_s[v]);
                }
                // The 2x2 solve (estimate_foreground_ml.py:158-175).
                S da[4];
#pragma unroll
                for (int d = 0; d < 4; ++d) da[d] = sv[d * kFgPad];
                const double b00 = st.dv[bf][j][0][lane];
                const double b01 = st.dv[bf][j][1][lane];
                const double b11 = st.dv[bf][j][2][lane];
                S res[2];
                fg_solve_channel<S>(da, sv[(4 + ch) * kFgPad], sv[(7 + ch) * kFgPad], b00, b01,
                                    b11, left, right, up, down, res);
                if (step_ok) {
                    st.out[bf][j][qf][lane] = res[0];
                    st.out[bf][j][qb][lane] = res[1];
                }
                if (active) {
                    if (lane == last_lane && r + 1 < h) {
                        // The strip's last row, tagged, for strip s+1.
                        auto* dst = halo + ((long long)s * w + xc) * 6;
                        __hip_atomic_store(dst + qf, fg_tag<S>(res[0], tag), __ATOMIC_RELAXED,
                                           __HIP_MEMORY_SCOPE_AGENT);
                        __hip_atomic_store(dst + qb, fg_tag<S>(res[1], tag), __ATOMIC_RELAXED,
                                           __HIP_MEMORY_SCOPE_AGENT);
                    }
                }
#pragma unroll
                for (int v = 0; v < 2; ++v) {
                    prev_res[v] = step_ok ? res[v] : prev_res[v];
                    cur_old[v] = step_ok ? next_old[v] : cur_old[v];
                }
            }
        } else {
            if (k > 0) fg_writeback<S>(st, bf ^ 1, t0 - C, s, lt, F, B, h, w);
            if (k + 1 < nchunks) {
                ok = fg_stage_rows<S>(st, bf ^ 1, t0 + C, s, lt, F, B, h, w, last_lane, halo,
                                      tag, waited) && ok;
                fg_store_lanes<S>(st, bf ^ 1, lt, lanes);
                if (k + 2 < nchunks)
                    fg_load_lanes<S>(lanes, t0 + 2 * C, s, lt, cs, cd, F, B, h, w);
            }
        }
        __syncthreads();
    }
    if (!computer) fg_writeback<S>(st, (nchunks - 1) & 1, (nchunks - 1) * C, s, lt, F, B, h, w);
    if (!ok && lt == 0) atomicExch(fail, 1);
}

// A whole level in one workgroup, when its colours and per-pixel
// coefficients fit in LDS: the level's coefficients (as k_fg_coef) and all of
// its n_iter sweeps in one launch, the sweeps pipelined. Sweep k updates
// pixel (y, x) at step x + y + 2k. Its left and upper neighbours then hold
// sweep k's values and its right and lower neighbours sweep k-1's, each
// written at the step before, which is what the reference's in-place raster
// sweeps read. A neighbour lies on the next or previous anti-diagonal, whose
// pixels are written only at steps of the other parity, so no value is read
// and written in the same step: one barrier per step orders everything.
constexpr int kFgSmallThreads = 1024;
constexpr size_t kFgSmallLds = 156 * 1024;
constexpr int kFgSmallMaxIter = 4096;  // sweeps per launch (more go to the strip kernel)

template <typename S>
constexpr size_t fg_small_bytes_per_px() {
    return sizeof(double) * kFgD + sizeof(S) * (10 + 6);  // b00..b11, da/b0/b1, F/B
}

template <typename S>
__global__ __launch_bounds__(kFgSmallThreads) void k_fg_small(const S* __restrict__ img,
                                                              const S* __restrict__ alpha,
                                                              int h, int w, S reg, S gw, S* F,
                                                              S* B, int n_iter) {
    extern __shared__ __align__(16) unsigned char fg_lds[];
    const int n = h * w;
    double* cd = reinterpret_cast<double*>(fg_lds);  // [3][n]: b00, b01, b11
    S* cs = reinterpret_cast<S*>(cd + kFgD * n);     // [10][n]: da l/r/u/d, b0 rgb, b1 rgb
    S* fb = cs + 10 * n;                             // [6][n]: F rgb, B rgb
    for (int i = threadIdx.x; i < n; i += kFgSmallThreads) {
        const int y = i / w, x = i - y * w;
        const int row = y * w;
        S an[4];
        an[0] = alpha[row + (x > 0 ? x - 1 : 0)];
        an[1] = alpha[row + (x + 1 < w ? x + 1 : x)];
        an[2] = alpha[(y > 0 ? y - 1 : 0) * w + x];
        an[3] = alpha[(y + 1 < h ? y + 1 : y) * w + x];
        const S I[3] = {img[i * 3], img[i * 3 + 1], img[i * 3 + 2]};
        PixelCoef<S> pc;
        pixel_coef<S>(alpha[i], an, I, reg, gw, pc);
#pragma unroll
        for (int d = 0; d < 4; ++d) cs[d * n + i] = pc.da[d];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            cs[(4 + c) * n + i] = pc.b0[c];
            cs[(7 + c) * n + i] = pc.b1[c];
            fb[c * n + i] = F[i * 3 + c];
            fb[(3 + c) * n + i] = B[i * 3 + c];
        }
        cd[i] = pc.b00;
        cd[n + i] = pc.b01;
        cd[2 * n + i] = pc.b11;
    }
    __syncthreads();
    const int items = n_iter * h;  // (sweep, row) pairs
    const int steps = (w - 1) + (h - 1) + 2 * (n_iter - 1) + 1;
    for (int t = 0; t < steps; ++t) {
        for (int j = threadIdx.x; j < items; j += kFgSmallThreads) {
            const int k = j / h, y = j - k * h;
            const int x = t - 2 * k - y;
            if (x < 0 || x >= w) continue;
            const int i = y * w + x;
            // Neighbours clamped to the pixel itself at the borders
            // (estimate_foreground_ml.py:144-157): its value before this sweep.
            const int il = x > 0 ? i - 1 : i, ir = x + 1 < w ? i + 1 : i;
            const int iu = y > 0 ? i - w : i, idn = y + 1 < h ? i + w : i;
            S left[6], right[6], up[6], down[6];
#pragma unroll
            for (int q = 0; q < 6; ++q) {
                left[q] = fb[q * n + il];
                right[q] = fb[q * n + ir];
                up[q] = fb[q * n + iu];
                down[q] = fb[q * n + idn];
            }
            S da[4], b0c[3], b1c[3];
#pragma unroll
            for (int d = 0; d < 4; ++d) da[d] = cs[d * n + i];
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                b0c[c] = cs[(4 + c) * n + i];
                b1c[c] = cs[(7 + c) * n + i];
            }
            S res[6];
            fg_solve_pixel<S>(da, b0c, b1c, cd[i], cd[n + i], cd[2 * n + i], left, right, up,
                              down, res);
#pragma unroll
            for (int q = 0; q < 6; ++q) fb[q * n + i] = res[q];
        }
        __syncthreads();
    }
    for (int i = threadIdx.x; i < n; i += kFgSmallThreads) {
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            F[i * 3 + c] = fb[c * n + i];
            B[i * 3 + c] = fb[(3 + c) * n + i];
        }
    }
}

// Nearest-neighbour resizes of one level (estimate_foreground_ml.py:5-55,
// 107-117): image/alpha from the full-resolution input, F/B from the previous
// level.
template <typename S>
__global__ void k_fg_resize(const S* __restrict__ img0, const S* __restrict__ alpha0, int h0,
                            int w0, const S* __restrict__ Fp, const S* __restrict__ Bp, int hp,
                            int wp, S* __restrict__ img, S* __restrict__ alpha,
                            S* __restrict__ F, S* __restrict__ B, int h, int w, int copy_input) {
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long long)h * w) return;
    const int y = (int)(i / w), x = (int)(i - (long long)y * w);
    if (copy_input) {
        int ys = (int)((long long)y * h0 / h), xs = (int)((long long)x * w0 / w);
        ys = ys < h0 - 1 ? ys : h0 - 1;
        xs = xs < w0 - 1 ? xs : w0 - 1;
        const long long q = (long long)ys * w0 + xs;
        img[i * 3] = img0[q * 3];
        img[i * 3 + 1] = img0[q * 3 + 1];
        img[i * 3 + 2] = img0[q * 3 + 2];
        alpha[i] = alpha0[q];
    }
    int yp = (int)((long long)y * hp / h), xp = (int)((long long)x * wp / w);
    yp = yp < hp - 1 ? yp : hp - 1;
    xp = xp < wp - 1 ? xp : wp - 1;
    const long long q = (long long)yp * wp + xp;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        F[i * 3 + c] = Fp[q * 3 + c];
        B[i * 3 + c] = Bp[q * 3 + c];
    }
}

// Sums of colours over alpha > 0.9 and alpha < 0.1 and the two counts
// (estimate_foreground_ml.py:78-95), in float64, deterministic.
template <typename S>
__global__ __launch_bounds__(kBlockMean) void k_fg_means(const S* __restrict__ img,
                                                         const S* __restrict__ alpha, long long n,
                                                         Reducer red, S* __restrict__ means) {
    double v[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (long long i = (long long)blockIdx.x * kBlockMean + threadIdx.x; i < n;
         i += (long long)gridDim.x * kBlockMean) {
        const double a = (double)alpha[i];
        if (a > 0.9) {
            v[0] += (double)img[i * 3];
            v[1] += (double)img[i * 3 + 1];
            v[2] += (double)img[i * 3 + 2];
            v[3] += 1.0;
        }
        if (a < 0.1) {
            v[4] += (double)img[i * 3];
            v[5] += (double)img[i * 3 + 1];
            v[6] += (double)img[i * 3 + 2];
            v[7] += 1.0;
        }
    }
    block_partials<kBlockMean, 8>(v, red);
}

// Mean colours from the partial sums, divided by count + 1e-5 in float64.
template <typename S>
__global__ __launch_bounds__(kBlockMean) void k_fg_means_final(Reducer red, unsigned nblocks,
                                                               S* __restrict__ means) {
    double acc[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        acc[k] = 0.0;
        for (unsigned i = threadIdx.x; i < nblocks; i += kBlockMean)
            acc[k] += red.partials[k * red.max_blocks + i];
    }
    block_sum<kBlockMean, 8>(acc);
    if (threadIdx.x == 0) {
        for (int c = 0; c < 3; ++c) {
            means[c] = (S)(acc[c] / (acc[3] + 1e-5));
            means[3 + c] = (S)(acc[4 + c] / (acc[7] + 1e-5));
        }
    }
}

template <typename S>
__global__ void k_fg_init(const S* __restrict__ means, S* __restrict__ F, S* __restrict__ B) {
    if (threadIdx.x < 3) {
        F[threadIdx.x] = means[threadIdx.x];
        B[threadIdx.x] = means[3 + threadIdx.x];
    }
}

template <typename W>
__global__ void k_fill_words(W* p, long long n) {
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = W(0);
}

__global__ void k_fill_i32(int* p, long long n) {
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

}  // namespace

// Device buffers of one foreground estimate, carved out of one workspace in
// a fixed order (each region 256-byte aligned). Sizes follow the pyramid:
// the ping-pong colour buffers and the resized inputs hold the levels below
// the finest, the coefficient planes and halo words the levels swept by the
// strip kernel.
struct FgLayout {
    enum { kImg, kAlpha, kFa, kBa, kFb, kBb, kMeans, kHalo, kFail, kCs, kCd, kRed, kCount };
    size_t off[kCount];
    size_t total;
    unsigned mean_blocks;
};

inline size_t fg_align(size_t b) { return (b + 255) / 256 * 256; }

template <typename S>
FgLayout fg_layout(int h0, int w0, int n_levels, const int* lh, const int* lw, int n_small,
                   int n_big, int small_size) {
    const long long n0 = (long long)h0 * w0;
    long long below = 1, strip = 0, halo = 0;  // points of the largest such level
    for (int l = 0; l <= n_levels; ++l) {
        const long long np = (long long)lh[l] * lw[l];
        if (l < n_levels && np > below) below = np;
        const int n_iter = (lw[l] <= small_size && lh[l] <= small_size) ? n_small : n_big;
        const bool one_wg = (size_t)np * fg_small_bytes_per_px<S>() <= kFgSmallLds &&
                            n_iter <= kFgSmallMaxIter;
        if (n_iter > 0 && !one_wg) {
            if (np > strip) strip = np;
            const long long nh = 6LL * div_up(lh[l], kWave) * lw[l];
            if (nh > halo) halo = nh;
        }
    }
    FgLayout L{};
    L.mean_blocks = div_up(n0, (long long)kBlockMean * 16) > 4096
                        ? 4096u
                        : div_up(n0, (long long)kBlockMean * 16);
    const size_t bytes[FgLayout::kCount] = {
        sizeof(S) * below * 3, sizeof(S) * below, sizeof(S) * below * 3, sizeof(S) * below * 3,
        sizeof(S) * below * 3, sizeof(S) * below * 3, sizeof(S) * 8,
        sizeof(typename FgWord<S>::U) * (size_t)(halo > 0 ? halo : 1), sizeof(int),
        sizeof(S) * (size_t)(strip > 0 ? strip : 1) * 10,
        sizeof(double) * (size_t)(strip > 0 ? strip : 1) * 3,
        sizeof(double) * 8 * ((size_t)L.mean_blocks + 1)};
    size_t at = 0;
    for (int k = 0; k < FgLayout::kCount; ++k) {
        L.off[k] = at;
        at += fg_align(bytes[k]);
    }
    L.total = at;
    return L;
}

// The pyramid of estimate_foreground_ml.py:98-183 on the buffers of `work`
// (laid out by fg_layout; its contents on entry do not matter: everything
// that must start at zero is cleared by a kernel).
template <typename S>
int fb_ml_run(const S* img0, const S* alpha0, int h0, int w0, int n_levels, const int* lh,
              const int* lw, double reg, int n_small, int n_big, int small_size, double gw,
              S* F_out, S* B_out, const FgLayout& lay, unsigned char* work, hipStream_t s) {
    const long long n0 = (long long)h0 * w0;
    auto at = [&](int k) { return static_cast<void*>(work + lay.off[k]); };
    S* imgb = static_cast<S*>(at(FgLayout::kImg));
    S* alphab = static_cast<S*>(at(FgLayout::kAlpha));
    S* Fa = static_cast<S*>(at(FgLayout::kFa));
    S* Ba = static_cast<S*>(at(FgLayout::kBa));
    S* Fb = static_cast<S*>(at(FgLayout::kFb));
    S* Bb = static_cast<S*>(at(FgLayout::kBb));
    S* means = static_cast<S*>(at(FgLayout::kMeans));
    int* fail = static_cast<int*>(at(FgLayout::kFail));
    S* cs = static_cast<S*>(at(FgLayout::kCs));
    double* cd = static_cast<double*>(at(FgLayout::kCd));
    Reducer red{static_cast<double*>(at(FgLayout::kRed)), (int)lay.mean_blocks + 1, 0};
    const unsigned mean_blocks = lay.mean_blocks;
    int rc = HM_OK;
    k_fill_i32<<<1, 64, 0, s>>>(fail, 1);
    hipError_t e = hipSuccess;
    k_fg_means<S><<<mean_blocks, kBlockMean, 0, s>>>(img0, alpha0, n0, red, means);
    k_fg_means_final<S><<<1, kBlockMean, 0, s>>>(red, mean_blocks, means);
    // Level "-1": 1x1 F/B initialised with the mean colours.
    S* Fp = Fa;
    S* Bp = Ba;
    k_fg_init<S><<<1, 64, 0, s>>>(means, Fp, Bp);
    int hp = 1, wp = 1;
    using W = typename FgWord<S>::U;
    W* halo = static_cast<W*>(at(FgLayout::kHalo));
    for (int l = 0; l <= n_levels; ++l) {
        const int h = lh[l], w = lw[l];
        const bool full = (h == h0 && w == w0);
        const bool last = (l == n_levels);
        S* Fc = last ? F_out : (Fp == Fa ? Fb : Fa);
        S* Bc = last ? B_out : (Bp == Ba ? Bb : Ba);
        const S* img = full ? img0 : imgb;
        const S* alpha = full ? alpha0 : alphab;
        k_fg_resize<S><<<div_up((long long)h * w, 256), 256, 0, s>>>(
            img0, alpha0, h0, w0, Fp, Bp, hp, wp, imgb, alphab, Fc, Bc, h, w, full ? 0 : 1);
        const int n_iter = (w <= small_size && h <= small_size) ? n_small : n_big;
        const size_t small_lds = (size_t)h * w * fg_small_bytes_per_px<S>();
        const bool one_wg = small_lds <= kFgSmallLds && n_iter <= kFgSmallMaxIter;
        if (n_iter > 0 && one_wg) {
            // The whole level in one workgroup (k_fg_small).
            rc = set_max_dynamic_lds(reinterpret_cast<const void*>(&k_fg_small<S>),
                                     (int)kFgSmallLds);
            if (rc != HM_OK) return rc;
            ProfScope prof(kProfFgSweep, (long long)h * w * n_iter, sizeof(S) == 4, s);
            k_fg_small<S><<<1, kFgSmallThreads, small_lds, s>>>(img, alpha, h, w, (S)reg, (S)gw,
                                                               Fc, Bc, n_iter);
        } else if (n_iter > 0) {
            k_fg_coef<S><<<div_up((long long)h * w, 256), 256, 0, s>>>(img, alpha, h, w, (S)reg,
                                                                      (S)gw, cs, cd);
            // Halo words of this level start untagged; sweeps cycle tags 1..3.
            const long long nh = 6LL * div_up(h, kWave) * w;
            k_fill_words<W><<<div_up(nh, 256), 256, 0, s>>>(halo, nh);
        }
        for (int it = 0; it < n_iter && !one_wg; ++it) {
            ProfScope prof(kProfFgSweep, (long long)h * w, sizeof(S) == 4, s);
            k_fg_sweep<S><<<div_up(h, kWave), kFgThreads, 0, s>>>(
                cs, cd, Fc, Bc, h, w, halo, 1u + (unsigned)(it % 3), fail);
        }
        e = hipGetLastError();
        if (e != hipSuccess) {
            set_error("foreground launch: %s", hipGetErrorString(e));
            return HM_ERR_HIP;
        }
        Fp = Fc;
        Bp = Bc;
        hp = h;
        wp = w;
    }
    int host_fail = 0;
    e = hipMemcpyAsync(&host_fail, fail, sizeof(int), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) {
        set_error("foreground: %s", hipGetErrorString(e));
        return HM_ERR_HIP;
    }
    if (host_fail) {
        set_error("foreground sweep: a strip timed out waiting for the strip above");
        return HM_ERR_HIP;
    }
    return HM_OK;
}

template <typename S>
int fb_ml_t(const S* img0, const S* alpha0, int h0, int w0, int n_levels, const int* lh,
            const int* lw, double reg, int n_small, int n_big, int small_size, double gw,
            S* F_out, S* B_out, hipStream_t s) {
    const FgLayout lay = fg_layout<S>(h0, w0, n_levels, lh, lw, n_small, n_big, small_size);
    void* work = nullptr;
    HM_TRY(dev_alloc(&work, lay.total));
    const int rc = fb_ml_run<S>(img0, alpha0, h0, w0, n_levels, lh, lw, reg, n_small, n_big,
                                small_size, gw, F_out, B_out, lay,
                                static_cast<unsigned char*>(work), s);
    dev_free(work);
    return rc;
}

// ------------------------------------------- the allocation-free variant
//
// hm_fb_ml_noalloc runs the same kernels on a workspace that is a file-scope
// device array, so it makes no allocation (for callers that cannot allocate:
// code under graph capture, serving loops, harnesses that forbid hipMalloc).
// Each device holds its own copy of the array for the life of the process;
// its address on a device is looked up once, on that device's first call.
// Every call on one device shares that device's array: calls that overlap
// on one device (on different streams or threads) race and must not be made.
constexpr size_t kFgScratchBytes = size_t(64) << 20;
static __device__ __align__(256) unsigned char scratch[kFgScratchBytes];

namespace {
constexpr int kMaxScratchDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxScratchDevices] = {};
}  // namespace

// The current device's copy of `scratch` (a device address usable in kernel
// arguments; the array's name itself is a host-side stand-in).
static int scratch_address(unsigned char** out) {
    int dev = 0;
    HM_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxScratchDevices) {
        set_error("hm_fb_ml_noalloc: device %d is beyond the %d devices served", dev,
                  kMaxScratchDevices);
        return HM_ERR_UNSUPPORTED;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        HM_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        g_scratch_ptr[dev] = p;
    }
    *out = static_cast<unsigned char*>(g_scratch_ptr[dev]);
    return HM_OK;
}

template <typename S>
int fb_ml_noalloc_t(const S* img0, const S* alpha0, int h0, int w0, int n_levels, const int* lh,
                    const int* lw, double reg, int n_small, int n_big, int small_size, double gw,
                    S* F_out, S* B_out, hipStream_t s) {
    const FgLayout lay = fg_layout<S>(h0, w0, n_levels, lh, lw, n_small, n_big, small_size);
    if (lay.total > kFgScratchBytes) {
        set_error("hm_fb_ml_noalloc: a %dx%d %s image needs %zu bytes of workspace, more than "
                  "the %zu bytes of the library's static workspace",
                  h0, w0, sizeof(S) == 4 ? "float32" : "float64", lay.total, kFgScratchBytes);
        return HM_ERR_ARG;
    }
    unsigned char* work = nullptr;
    HM_TRY(scratch_address(&work));
    return fb_ml_run<S>(img0, alpha0, h0, w0, n_levels, lh, lw, reg, n_small, n_big, small_size,
                        gw, F_out, B_out, lay, work, s);
}

}  // namespace hm

static int fb_ml_check(const void* image, const void* alpha, int h0, int w0, int n_levels,
                       const int* level_h, const int* level_w, const void* F, const void* B,
                       const char* who) {
    using namespace hm;
    if (!image || !alpha || !F || !B || !level_h || !level_w || h0 < 1 || w0 < 1 ||
        n_levels < 1) {
        set_error("%s: invalid argument", who);
        return HM_ERR_ARG;
    }
    if (level_h[n_levels] != h0 || level_w[n_levels] != w0 || level_h[0] != 1 ||
        level_w[0] != 1) {
        set_error("%s: the pyramid must run from 1x1 to the image size", who);
        return HM_ERR_ARG;
    }
    for (int l = 0; l <= n_levels; ++l)
        if (level_h[l] < 1 || level_w[l] < 1 || level_h[l] > h0 || level_w[l] > w0) {
            set_error("%s: invalid level size", who);
            return HM_ERR_ARG;
        }
    return HM_OK;
}

extern "C" int hm_fb_ml(const void* image, const void* alpha, int h0, int w0, int n_levels,
                        const int* level_h, const int* level_w, double regularization,
                        int n_small, int n_big, int small_size, double gradient_weight, int prec,
                        void* F, void* B, void* stream) {
    using namespace hm;
    HM_TRY(fb_ml_check(image, alpha, h0, w0, n_levels, level_h, level_w, F, B, "hm_fb_ml"));
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (prec == HM_F32)
        return fb_ml_t<float>(static_cast<const float*>(image), static_cast<const float*>(alpha),
                              h0, w0, n_levels, level_h, level_w, regularization, n_small, n_big,
                              small_size, gradient_weight, static_cast<float*>(F),
                              static_cast<float*>(B), s);
    return fb_ml_t<double>(static_cast<const double*>(image), static_cast<const double*>(alpha),
                           h0, w0, n_levels, level_h, level_w, regularization, n_small, n_big,
                           small_size, gradient_weight, static_cast<double*>(F),
                           static_cast<double*>(B), s);
}

extern "C" int hm_fb_ml_noalloc(const void* image, const void* alpha, int h0, int w0,
                                int n_levels, const int* level_h, const int* level_w,
                                double regularization, int n_small, int n_big, int small_size,
                                double gradient_weight, int prec, void* F, void* B,
                                void* stream) {
    using namespace hm;
    HM_TRY(fb_ml_check(image, alpha, h0, w0, n_levels, level_h, level_w, F, B,
                       "hm_fb_ml_noalloc"));
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (prec == HM_F32)
        return fb_ml_noalloc_t<float>(static_cast<const float*>(image),
                                      static_cast<const float*>(alpha), h0, w0, n_levels,
                                      level_h, level_w, regularization, n_small, n_big,
                                      small_size, gradient_weight, static_cast<float*>(F),
                                      static_cast<float*>(B), s);
    return fb_ml_noalloc_t<double>(static_cast<const double*>(image),
                                   static_cast<const double*>(alpha), h0, w0, n_levels, level_h,
                                   level_w, regularization, n_small, n_big, small_size,
                                   gradient_weight, static_cast<double*>(F),
                                   static_cast<double*>(B), s);
}
