// Multilevel foreground/background estimation on gfx950
// (pymatting/foreground/estimate_foreground_ml.py:61-183).
//
// The reference updates pixels in place in raster order (Gauss-Seidel): the
// new value of (y, x) uses the NEW values of (y, x-1) and (y-1, x) and the
// OLD values of (y, x+1) and (y+1, x). This file reproduces that order
// exactly with a strip-pipelined wavefront:
//   - one 64-lane wavefront owns a strip of 64 rows; lane y owns row y and at
//     step t updates column x = t - y, so lane y runs one column behind lane
//     y-1. The new value of the pixel above is lane y-1's result of the
//     previous step, and the old value of the pixel below is the "next old"
//     register lane y+1 loaded for its right neighbour; both move by
//     cross-lane shuffles, never through memory.
//   - strip s reads the new last row of strip s-1 from global memory after
//     strip s-1 has published (device-scope release/acquire) that the column
//     is final; strip s-1 reads the old first row of strip s before strip s
//     may overwrite it, which the same publish order guarantees.
// Waits are bounded: a strip that waits too long raises a failure flag and
// finishes, so every wave exits.
//
// Arithmetic follows the types Numba gives the reference: with float32
// arrays, a0, a00, da and the b vector are float32 while a1 = 1.0 - a0 and
// everything derived from it are float64. Compiled with -ffp-contract=off.
#include "common.h"
#include "internal.h"

namespace hm {

namespace {

constexpr int kFgChunk = 8;          // steps between progress publications
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

template <typename S>
__device__ __forceinline__ S shfl_up1(S v) { return __shfl_up(v, 1, kWave); }
template <typename S>
__device__ __forceinline__ S shfl_down1(S v) { return __shfl_down(v, 1, kWave); }

template <typename S>
__device__ __forceinline__ S load_acquire_val(const S* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// One in-place Gauss-Seidel sweep over an (h, w) level. img (h,w,3), alpha
// (h,w), F/B (h,w,3). progress[s] holds (tag << 32) | columns finished by the
// last row of strip s; tag identifies this sweep.
template <typename S>
__global__ __launch_bounds__(64) void k_fg_sweep(const S* __restrict__ img,
                                                 const S* __restrict__ alpha, S* F, S* B, int h,
                                                 int w, S reg, S gw,
                                                 unsigned long long* progress,
                                                 unsigned long long tag, int* fail) {
    const int s = blockIdx.x;
    const int lane = threadIdx.x;
    const int r = s * kWave + lane;
    const bool row_ok = r < h;
    const int rc = row_ok ? r : h - 1;
    const long long row = (long long)rc * w;
    const int last_lane = min(kWave, h - s * kWave) - 1;  // lane owning the strip's last row

    S cur_old[6], next_old[6], prev_res[6];
    // Old values at column 0 (the first step's "current" pixel).
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        cur_old[c] = F[row * 3 + c];
        cur_old[3 + c] = B[row * 3 + c];
        prev_res[c] = cur_old[c];
        prev_res[3 + c] = cur_old[3 + c];
    }
    const int steps = w + kWave - 1;
    long long waited = 0;
    bool failed = false;
    for (int t0 = 0; t0 < steps; t0 += kFgChunk) {
        // Strip s needs row 64s-1 finished up to the columns lane 0 touches.
        if (s > 0) {
            const unsigned long long need = (unsigned long long)min(t0 + kFgChunk, w);
            while (!failed) {
                const unsigned long long v = __hip_atomic_load(&progress[s - 1], __ATOMIC_ACQUIRE,
                                                               __HIP_MEMORY_SCOPE_AGENT);
                if ((v >> 32) == tag && (v & 0xffffffffull) >= need) break;
                if (++waited > kSpinLimit) {
                    failed = true;
                    break;
                }
                __builtin_amdgcn_s_sleep(1);
            }
        }
        const int t1 = min(t0 + kFgChunk, steps);
        for (int t = t0; t < t1; ++t) {
            const int x = t - lane;
            const bool active = row_ok && x >= 0 && x < w;
            const int xc = x < 0 ? 0 : (x >= w ? w - 1 : x);
            // Old values one column to the right (kept for the next step and
            // for lane-1's down neighbour).
            const int xn = xc + 1 < w ? xc + 1 : xc;
            const long long pn = (row + (x + 1 >= 0 && x + 1 < w ? x + 1 : xn)) * 3;
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                next_old[c] = F[pn + c];
                next_old[3 + c] = B[pn + c];
            }
            // Neighbour values.
            S left[6], right[6], up[6], down[6];
            S from_above[6], from_below[6];
#pragma unroll
            for (int k = 0; k < 6; ++k) {
                from_above[k] = shfl_up1(prev_res[k]);    // lane-1's result at this column
                from_below[k] = shfl_down1(next_old[k]);  // lane+1's old value at this column
            }
            const long long p = (row + xc) * 3;
#pragma unroll
            for (int k = 0; k < 6; ++k) {
                left[k] = x > 0 ? prev_res[k] : cur_old[k];
                right[k] = x + 1 < w ? next_old[k] : cur_old[k];
            }
            if (r == 0) {
#pragma unroll
                for (int k = 0; k < 6; ++k) up[k] = cur_old[k];
            } else if (lane > 0) {
#pragma unroll
                for (int k = 0; k < 6; ++k) up[k] = from_above[k];
            } else {
                const long long pu = ((long long)(r - 1) * w + xc) * 3;
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    up[c] = load_acquire_val(F + pu + c);
                    up[3 + c] = load_acquire_val(B + pu + c);
                }
            }
            if (r >= h - 1) {
#pragma unroll
                for (int k = 0; k < 6; ++k) down[k] = cur_old[k];
            } else if (lane < last_lane) {
#pragma unroll
                for (int k = 0; k < 6; ++k) down[k] = from_below[k];
            } else {
                const long long pd = ((long long)(r + 1) * w + xc) * 3;
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    down[c] = load_acquire_val(F + pd + c);
                    down[3 + c] = load_acquire_val(B + pd + c);
                }
            }
            // The 2x2 solve.
            const long long pa = row + xc;
            const S a0 = alpha[pa];
            S an[4];
            an[0] = alpha[row + (xc > 0 ? xc - 1 : 0)];
            an[1] = alpha[row + (xc + 1 < w ? xc + 1 : xc)];
            an[2] = alpha[(long long)(rc > 0 ? rc - 1 : 0) * w + xc];
            an[3] = alpha[(long long)(rc + 1 < h ? rc + 1 : rc) * w + xc];
            S I[3] = {img[p], img[p + 1], img[p + 2]};
            PixelCoef<S> pc;
            pixel_coef<S>(a0, an, I, reg, gw, pc);
            S res[6];
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                S b0 = pc.b0[c], b1 = pc.b1[c];
                b0 += pc.da[0] * left[c];
                b1 += pc.da[0] * left[3 + c];
                b0 += pc.da[1] * right[c];
                b1 += pc.da[1] * right[3 + c];
                b0 += pc.da[2] * up[c];
                b1 += pc.da[2] * up[3 + c];
                b0 += pc.da[3] * down[c];
                b1 += pc.da[3] * down[3 + c];
                res[c] = clamp01_store<S>(pc.b00 * (double)b0 + pc.b01 * (double)b1);
                res[3 + c] = clamp01_store<S>(pc.b01 * (double)b0 + pc.b11 * (double)b1);
            }
            if (active) {
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    F[p + c] = res[c];
                    B[p + c] = res[3 + c];
                }
            }
#pragma unroll
            for (int k = 0; k < 6; ++k) {
                prev_res[k] = res[k];
                cur_old[k] = next_old[k];
            }
        }
        // Publish how many columns the strip's last row has finished.
        if (lane == last_lane) {
            const int done_cols = min(max(t1 - last_lane, 0), w);
            __threadfence();
            __hip_atomic_store(&progress[s], (tag << 32) | (unsigned long long)done_cols,
                               __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
    if (failed && lane == 0) atomicExch(fail, 1);
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

__global__ void k_fill_u64(unsigned long long* p, long long n) {
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0ull;
}

}  // namespace

template <typename S>
int fb_ml_t(const S* img0, const S* alpha0, int h0, int w0, int n_levels, const int* lh,
            const int* lw, double reg, int n_small, int n_big, int small_size, double gw,
            S* F_out, S* B_out, hipStream_t s) {
    const long long n0 = (long long)h0 * w0;
    void *imgb = nullptr, *alphab = nullptr, *Fa = nullptr, *Ba = nullptr, *Fb = nullptr,
         *Bb = nullptr, *means = nullptr, *prog = nullptr, *fail = nullptr;
    Reducer red{};
    int rc = HM_OK;
    const int max_strips = (int)div_up(h0, kWave);
    auto cleanup = [&]() {
        dev_free(imgb); dev_free(alphab); dev_free(Fa); dev_free(Ba); dev_free(Fb); dev_free(Bb);
        dev_free(means); dev_free(prog); dev_free(fail);
        free_reducer(&red);
    };
    const unsigned mean_blocks = div_up(n0, (long long)kBlockMean * 16) > 4096
                                     ? 4096u
                                     : div_up(n0, (long long)kBlockMean * 16);
    if ((rc = dev_alloc(&imgb, sizeof(S) * n0 * 3)) != HM_OK ||
        (rc = dev_alloc(&alphab, sizeof(S) * n0)) != HM_OK ||
        (rc = dev_alloc(&Fa, sizeof(S) * n0 * 3)) != HM_OK ||
        (rc = dev_alloc(&Ba, sizeof(S) * n0 * 3)) != HM_OK ||
        (rc = dev_alloc(&Fb, sizeof(S) * n0 * 3)) != HM_OK ||
        (rc = dev_alloc(&Bb, sizeof(S) * n0 * 3)) != HM_OK ||
        (rc = dev_alloc(&means, sizeof(S) * 8)) != HM_OK ||
        (rc = dev_alloc(&prog, sizeof(unsigned long long) * (max_strips + 1))) != HM_OK ||
        (rc = dev_alloc(&fail, sizeof(int))) != HM_OK ||
        (rc = make_reducer(&red, (int)mean_blocks + 1, 8)) != HM_OK) {
        cleanup();
        return rc;
    }
    k_fill_u64<<<div_up(max_strips + 1, 64), 64, 0, s>>>(static_cast<unsigned long long*>(prog),
                                                         max_strips + 1);
    hipError_t e = hipMemsetAsync(fail, 0, sizeof(int), s);
    if (e != hipSuccess) { cleanup(); set_error("hipMemsetAsync: %s", hipGetErrorString(e)); return HM_ERR_HIP; }
    k_fg_means<S><<<mean_blocks, kBlockMean, 0, s>>>(img0, alpha0, n0, red, static_cast<S*>(means));
    k_fg_means_final<S><<<1, kBlockMean, 0, s>>>(red, mean_blocks, static_cast<S*>(means));
    // Level "-1": 1x1 F/B initialised with the mean colours.
    S* Fp = static_cast<S*>(Fa);
    S* Bp = static_cast<S*>(Ba);
    k_fg_init<S><<<1, 64, 0, s>>>(static_cast<S*>(means), Fp, Bp);
    int hp = 1, wp = 1;
    unsigned long long tag = 0;
    for (int l = 0; l <= n_levels; ++l) {
        const int h = lh[l], w = lw[l];
        const bool full = (h == h0 && w == w0);
        const bool last = (l == n_levels);
        S* Fc = last ? F_out : (Fp == static_cast<S*>(Fa) ? static_cast<S*>(Fb) : static_cast<S*>(Fa));
        S* Bc = last ? B_out : (Bp == static_cast<S*>(Ba) ? static_cast<S*>(Bb) : static_cast<S*>(Ba));
        const S* img = full ? img0 : static_cast<S*>(imgb);
        const S* alpha = full ? alpha0 : static_cast<S*>(alphab);
        k_fg_resize<S><<<div_up((long long)h * w, 256), 256, 0, s>>>(
            img0, alpha0, h0, w0, Fp, Bp, hp, wp, static_cast<S*>(imgb), static_cast<S*>(alphab),
            Fc, Bc, h, w, full ? 0 : 1);
        const int n_iter = (w <= small_size && h <= small_size) ? n_small : n_big;
        for (int it = 0; it < n_iter; ++it) {
            ++tag;
            ProfScope prof(kProfFgSweep, (long long)h * w, sizeof(S) == 4, s);
            k_fg_sweep<S><<<div_up(h, kWave), kWave, 0, s>>>(
                img, alpha, Fc, Bc, h, w, (S)reg, (S)gw, static_cast<unsigned long long*>(prog),
                tag, static_cast<int*>(fail));
        }
        e = hipGetLastError();
        if (e != hipSuccess) { cleanup(); set_error("foreground launch: %s", hipGetErrorString(e)); return HM_ERR_HIP; }
        Fp = Fc;
        Bp = Bc;
        hp = h;
        wp = w;
    }
    int host_fail = 0;
    e = hipMemcpyAsync(&host_fail, fail, sizeof(int), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    cleanup();
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

}  // namespace hm

extern "C" int hm_fb_ml(const void* image, const void* alpha, int h0, int w0, int n_levels,
                        const int* level_h, const int* level_w, double regularization,
                        int n_small, int n_big, int small_size, double gradient_weight, int prec,
                        void* F, void* B, void* stream) {
    using namespace hm;
    if (!image || !alpha || !F || !B || !level_h || !level_w || h0 < 1 || w0 < 1 ||
        n_levels < 1) {
        set_error("hm_fb_ml: invalid argument");
        return HM_ERR_ARG;
    }
    if (level_h[n_levels] != h0 || level_w[n_levels] != w0 || level_h[0] != 1 ||
        level_w[0] != 1) {
        set_error("hm_fb_ml: the pyramid must run from 1x1 to the image size");
        return HM_ERR_ARG;
    }
    for (int l = 0; l <= n_levels; ++l)
        if (level_h[l] < 1 || level_w[l] < 1 || level_h[l] > h0 || level_w[l] > w0) {
            set_error("hm_fb_ml: invalid level size");
            return HM_ERR_ARG;
        }
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
