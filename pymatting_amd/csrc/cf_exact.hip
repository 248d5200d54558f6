// Bit-exact assembly of the closed-form matting Laplacian on gfx950.
//
// Reproduces the entries of pymatting/laplacian/cf_laplacian.py:5-129 bit for
// bit. Each window's mean, covariance and cofactor inverse follow the
// reference's operation order, and every matrix entry is gathered (no atomics)
// as the sum over the windows containing both pixels in raster order of the
// window centre -- the order in which the reference's serial scatter loop adds
// them. This file is compiled with -ffp-contract=off so no multiply-add is
// fused, matching the reference's arithmetic.
#include "common.h"
#include "internal.h"

namespace hm {

__global__ void k_hwc_to_planar_f64(const double* __restrict__ hwc, double* __restrict__ out,
                                    long long n) {
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    out[t] = hwc[3 * t];
    out[n + t] = hwc[3 * t + 1];
    out[2 * n + t] = hwc[3 * t + 2];
}

__global__ void k_planar_to_f32(const double* __restrict__ in, float* __restrict__ out,
                                long long m) {
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < m) out[t] = (float)in[t];
}

__global__ __launch_bounds__(256) void k_trimap_classes(const double* __restrict__ trimap,
                                                        long long n, double fg_thr, double bg_thr,
                                                        uint8_t* __restrict__ cls,
                                                        unsigned long long* __restrict__ part) {
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    int fg = 0, bg = 0;
    if (t < n) {
        const double v = trimap[t];
        fg = v >= fg_thr;  // util.py:509
        bg = v <= bg_thr;  // util.py:510
        // A pixel can be both only if fg_thr <= bg_thr; the class byte keeps fg then.
        cls[t] = fg ? 1 : (bg ? 2 : 0);
    }
    __shared__ unsigned cnt[2];
    if (threadIdx.x == 0) cnt[0] = cnt[1] = 0;
    __syncthreads();
    const unsigned long long bf = __ballot(fg), bb = __ballot(bg);
    if ((threadIdx.x % kWave) == 0) {
        atomicAdd(&cnt[0], (unsigned)__popcll(bf));
        atomicAdd(&cnt[1], (unsigned)__popcll(bb));
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = cnt[0];
        part[2 * blockIdx.x + 1] = cnt[1];
    }
}

__global__ __launch_bounds__(256) void k_trimap_counts(const unsigned long long* __restrict__ part,
                                                       unsigned nblocks, int64_t* __restrict__ counts) {
    __shared__ unsigned long long s[2][256];
    unsigned long long f = 0, b = 0;
    for (unsigned i = threadIdx.x; i < nblocks; i += 256) {
        f += part[2 * i];
        b += part[2 * i + 1];
    }
    s[0][threadIdx.x] = f;
    s[1][threadIdx.x] = b;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) {
            s[0][threadIdx.x] += s[0][threadIdx.x + off];
            s[1][threadIdx.x] += s[1][threadIdx.x + off];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        counts[0] = (int64_t)s[0][0];
        counts[1] = (int64_t)s[1][0];
    }
}

// Per-window data, indexed by the window's centre pixel: mean (s/9) per
// channel, the six inverse-covariance entries, and a validity flag.
// Layout [10][n].
__global__ void k_window_exact(const double* __restrict__ I, int h, int w, double eps,
                               const uint8_t* __restrict__ cls, int skip_known,
                               double* __restrict__ wc) {
    const long long n = (long long)h * w;
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int y = (int)(t / w), x = (int)(t - (long long)y * w);
    double* flag = wc + 9 * n;
    if (x < 1 || x > w - 2 || y < 1 || y > h - 2) {
        flag[t] = 0.0;
        return;
    }
    if (skip_known && cls) {
        bool all_known = true;
        for (int dy = -1; dy <= 1; ++dy)
            for (int dx = -1; dx <= 1; ++dx) all_known &= cls[t + (long long)dy * w + dx] != 0;
        if (all_known) {
            flag[t] = 0.0;
            return;
        }
    }
    double mean[3];
    for (int c = 0; c < 3; ++c) {  // cf_laplacian.py:40-50
        double s = 0.0;
        for (int dy = -1; dy <= 1; ++dy)
            for (int dx = -1; dx <= 1; ++dx) s += I[c * n + t + (long long)dy * w + dx];
        mean[c] = s / 9;
    }
    double a00 = eps, a01 = 0.0, a02 = 0.0, a11 = eps, a12 = 0.0, a22 = eps;
    for (int dy = -1; dy <= 1; ++dy) {  // cf_laplacian.py:53-67
        for (int dx = -1; dx <= 1; ++dx) {
            const long long q = t + (long long)dy * w + dx;
            const double c0 = I[q] - mean[0], c1 = I[n + q] - mean[1], c2 = I[2 * n + q] - mean[2];
            a00 += c0 * c0;
            a01 += c0 * c1;
            a02 += c0 * c2;
            a11 += c1 * c1;
            a12 += c1 * c2;
            a22 += c2 * c2;
        }
    }
    a00 /= 9; a01 /= 9; a02 /= 9; a11 /= 9; a12 /= 9; a22 /= 9;
    // cf_laplacian.py:76-92: negated determinant and matching cofactors.
    const double det = a00 * a12 * a12 + a01 * a01 * a22 + a02 * a02 * a11 - a00 * a11 * a22 -
                       2 * a01 * a02 * a12;
    const double inv_det = 1.0 / det;
    wc[t] = mean[0];
    wc[n + t] = mean[1];
    wc[2 * n + t] = mean[2];
    wc[3 * n + t] = (a12 * a12 - a11 * a22) * inv_det;  // m00
    wc[4 * n + t] = (a01 * a22 - a02 * a12) * inv_det;  // m01
    wc[5 * n + t] = (a02 * a11 - a01 * a12) * inv_det;  // m02
    wc[6 * n + t] = (a02 * a02 - a00 * a22) * inv_det;  // m11
    wc[7 * n + t] = (a00 * a12 - a01 * a02) * inv_det;  // m12
    wc[8 * n + t] = (a01 * a01 - a00 * a11) * inv_det;  // m22
    flag[t] = 1.0;
}

// L_ij for j = i + (dy, dx), summed over windows in raster order
// (cf_laplacian.py:95-129). Requires j inside the image.
__device__ __forceinline__ double entry_exact(const double* __restrict__ I,
                                              const double* __restrict__ wc, int h, int w,
                                              int yi, int xi, int yj, int xj) {
    const long long n = (long long)h * w;
    const long long i = (long long)yi * w + xi, j = (long long)yj * w + xj;
    const double Ii0 = I[i], Ii1 = I[n + i], Ii2 = I[2 * n + i];
    const double Ij0 = I[j], Ij1 = I[n + j], Ij2 = I[2 * n + j];
    const double delta = (i == j) ? 1.0 : 0.0;
    double value = 0.0;
    const int ky0 = max(yi, yj) - 1, ky1 = min(yi, yj) + 1;
    const int kx0 = max(xi, xj) - 1, kx1 = min(xi, xj) + 1;
    for (int ky = ky0; ky <= ky1; ++ky) {
        if (ky < 1 || ky > h - 2) continue;
        for (int kx = kx0; kx <= kx1; ++kx) {
            if (kx < 1 || kx > w - 2) continue;
            const long long k = (long long)ky * w + kx;
            if (wc[9 * n + k] == 0.0) continue;
            const double mu0 = wc[k], mu1 = wc[n + k], mu2 = wc[2 * n + k];
            const double m00 = wc[3 * n + k], m01 = wc[4 * n + k], m02 = wc[5 * n + k];
            const double m11 = wc[6 * n + k], m12 = wc[7 * n + k], m22 = wc[8 * n + k];
            const double s = Ii0 - mu0, tt = Ii1 - mu1, u = Ii2 - mu2;
            const double c0 = m00 * s + m01 * tt + m02 * u;
            const double c1 = m01 * s + m11 * tt + m12 * u;
            const double c2 = m02 * s + m12 * tt + m22 * u;
            const double temp = c0 * (Ij0 - mu0) + c1 * (Ij1 - mu1) + c2 * (Ij2 - mu2);
            value += delta - (1 + temp) / 9;
        }
    }
    return value;
}

// Assembles slots s_lo..24 per pixel. Output element (i, s) lives at
// values[i * stride_i + (s - s_lo) * stride_s] (s_lo = 12: the 13 symmetric
// slots of the stored stencil form). mode: HM_OP_* adjustments of the
// stored operator (lambda on the diagonal of known rows, or the
// hard-constraint masking). indices/indptr (nullable) follow the reference's
// CSR layout.
__global__ void k_assemble_exact(const double* __restrict__ I, const double* __restrict__ wc,
                                 int h, int w, const uint8_t* __restrict__ cls, int mode,
                                 double lambda, double* __restrict__ values,
                                 long long stride_i, long long stride_s, int s_lo,
                                 int64_t* __restrict__ indices, int64_t* __restrict__ indptr) {
    const long long n = (long long)h * w;
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int yi = (int)(i / w), xi = (int)(i - (long long)yi * w);
    const bool known_i = cls && cls[i] != 0;
    for (int s = s_lo; s < 25; ++s) {
        const int yj = yi + s / 5 - 2, xj = xi + s % 5 - 2;
        double v = 0.0;
        int64_t col = 0;
        if (yj >= 0 && yj < h && xj >= 0 && xj < w) {
            col = (int64_t)xj + (int64_t)yj * w;
            const bool known_j = cls && cls[col] != 0;
            if (mode == HM_OP_MASKED && known_i) {
                v = (s == 12) ? 1.0 : 0.0;
            } else if (mode == HM_OP_MASKED && known_j) {
                v = 0.0;
            } else {
                v = entry_exact(I, wc, h, w, yi, xi, yj, xj);
                if (mode == HM_OP_LAMBDA && s == 12 && known_i) v = v + lambda;
            }
        }
        values[i * stride_i + (s - s_lo) * stride_s] = v;
        if (indices) indices[i * 25 + s] = col;
    }
    if (indptr) {
        indptr[i + 1] = (i + 1) * 25;
        if (i == 0) indptr[0] = 0;
    }
}

// The stored symmetric form (slots 12..24, planes [13][n]) with its float32
// copy, diagonal and inverse diagonal, per tile of kAsmX x kAsmY pixels. The
// window data (flag, mean, inverse covariance) of the tile's windows and the
// pixels of its 5 x 5 neighbourhoods are staged in LDS; each pixel i then
// forms c_k = M_k (I_i - mu_k) once for each of its nine windows and sums
// L_ij = sum_k [delta_ij - (1 + c_k . (I_j - mu_k)) / 9] over the windows
// containing both pixels in raster order of the centre, with the operations
// of entry_exact, so every entry is the same double.
constexpr int kAsmX = 32, kAsmY = 8;
constexpr int kAsmWX = kAsmX + 2, kAsmWY = kAsmY + 2;  // window centres
constexpr int kAsmIX = kAsmX + 4, kAsmIY = kAsmY + 4;  // pixels

__global__ __launch_bounds__(kAsmX * kAsmY) void k_assemble_sym(
    const double* __restrict__ I, const double* __restrict__ wc, int h, int w,
    const uint8_t* __restrict__ cls, int mode, double lambda, double* __restrict__ c64,
    float* __restrict__ c32, double* __restrict__ d64, float* __restrict__ d32,
    double* __restrict__ id64, float* __restrict__ id32) {
    __shared__ double sw[10][kAsmWY * kAsmWX];  // flag, mu0..2, m00 m01 m02 m11 m12 m22
    __shared__ double si[3][kAsmIY * kAsmIX];
    const long long n = (long long)h * w;
    const int x0 = blockIdx.x * kAsmX, y0 = blockIdx.y * kAsmY;
    const int tid = threadIdx.x;
    for (int t = tid; t < kAsmWY * kAsmWX; t += kAsmX * kAsmY) {
        const int ky = y0 - 1 + t / kAsmWX, kx = x0 - 1 + t % kAsmWX;
        const bool in = ky >= 0 && ky < h && kx >= 0 && kx < w;
        const long long k = in ? (long long)ky * w + kx : 0;
        sw[0][t] = in ? wc[9 * n + k] : 0.0;  // outside the image: no window
#pragma unroll
        for (int q = 0; q < 9; ++q) sw[1 + q][t] = in ? wc[q * n + k] : 0.0;
    }
    for (int t = tid; t < kAsmIY * kAsmIX; t += kAsmX * kAsmY) {
        const int y = y0 - 2 + t / kAsmIX, x = x0 - 2 + t % kAsmIX;
        const bool in = y >= 0 && y < h && x >= 0 && x < w;
        const long long q = in ? (long long)y * w + x : 0;
#pragma unroll
        for (int c = 0; c < 3; ++c) si[c][t] = in ? I[c * n + q] : 0.0;
    }
    __syncthreads();
    const int tx = tid % kAsmX, ty = tid / kAsmX;
    const int xi = x0 + tx, yi = y0 + ty;
    if (xi >= w || yi >= h) return;
    const long long i = (long long)yi * w + xi;
    const int ib = (ty + 2) * kAsmIX + tx + 2;
    const double Ii0 = si[0][ib], Ii1 = si[1][ib], Ii2 = si[2][ib];
    // Window (a, b) is centred at (yi + a - 1, xi + b - 1); windows outside
    // rows/columns 1..h-2 / 1..w-2 have flag 0 in wc. The windows are visited
    // in raster order of the centre, each adding its term to every stored
    // slot whose pixel j it also contains, so each slot's sum keeps the
    // order of entry_exact; slot (dy, dx) lies in windows
    // a = max(dy, 0) .. min(dy, 0) + 2 (b likewise).
    double acc[13];  // stored slots 12..24
#pragma unroll
    for (int s = 12; s < 25; ++s) acc[s - 12] = 0.0;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
#pragma unroll
        for (int b = 0; b < 3; ++b) {
            const int wi = (ty + a) * kAsmWX + tx + b;
            if (sw[0][wi] == 0.0) continue;
            const double mu0 = sw[1][wi], mu1 = sw[2][wi], mu2 = sw[3][wi];
            const double m00 = sw[4][wi], m01 = sw[5][wi], m02 = sw[6][wi];
            const double m11 = sw[7][wi], m12 = sw[8][wi], m22 = sw[9][wi];
            const double s0 = Ii0 - mu0, t0 = Ii1 - mu1, u0 = Ii2 - mu2;
            const double c0 = m00 * s0 + m01 * t0 + m02 * u0;
            const double c1 = m01 * s0 + m11 * t0 + m12 * u0;
            const double c2 = m02 * s0 + m12 * t0 + m22 * u0;
#pragma unroll
            for (int s = 12; s < 25; ++s) {
                const int dy = s / 5 - 2, dx = s % 5 - 2;
                if (a < (dy > 0 ? dy : 0) || a > (dy < 0 ? dy : 0) + 2) continue;
                if (b < (dx > 0 ? dx : 0) || b > (dx < 0 ? dx : 0) + 2) continue;
                const int jb = (ty + 2 + dy) * kAsmIX + tx + 2 + dx;
                const double temp = c0 * (si[0][jb] - mu0) + c1 * (si[1][jb] - mu1) +
                                    c2 * (si[2][jb] - mu2);
                const double delta = (s == 12) ? 1.0 : 0.0;
                acc[s - 12] += delta - (1 + temp) / 9;
            }
        }
    }
    const bool known_i = cls && cls[i] != 0;
#pragma unroll
    for (int s = 12; s < 25; ++s) {
        const int dy = s / 5 - 2, dx = s % 5 - 2;
        const int yj = yi + dy, xj = xi + dx;
        double v = 0.0;
        if (yj < h && xj >= 0 && xj < w) {  // yj >= yi >= 0
            const long long j = (long long)yj * w + xj;
            const bool known_j = cls && cls[j] != 0;
            if (mode == HM_OP_MASKED && known_i) {
                v = (s == 12) ? 1.0 : 0.0;
            } else if (mode == HM_OP_MASKED && known_j) {
                v = 0.0;
            } else {
                v = acc[s - 12];
                if (mode == HM_OP_LAMBDA && s == 12 && known_i) v = v + lambda;
            }
        }
        const long long o = (long long)(s - 12) * n + i;
        c64[o] = v;
        c32[o] = (float)v;
        if (s == 12) {
            d64[i] = v;
            d32[i] = (float)v;
            id64[i] = 1.0 / v;
            id32[i] = (float)(1.0 / v);
        }
    }
}

// Diagonal of the mode-adjusted operator (float64 and float32 copies).
__global__ void k_diag_exact(const double* __restrict__ I, const double* __restrict__ wc, int h,
                             int w, const uint8_t* __restrict__ cls, int mode, double lambda,
                             double* __restrict__ d64, float* __restrict__ d32,
                             double* __restrict__ id64, float* __restrict__ id32) {
    const long long n = (long long)h * w;
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int yi = (int)(i / w), xi = (int)(i - (long long)yi * w);
    const bool known_i = cls && cls[i] != 0;
    double v;
    if (mode == HM_OP_MASKED && known_i) {
        v = 1.0;
    } else {
        v = entry_exact(I, wc, h, w, yi, xi, yi, xi);
        if (mode == HM_OP_LAMBDA && known_i) v = v + lambda;
    }
    d64[i] = v;
    if (d32) d32[i] = (float)v;
    if (id64) id64[i] = 1.0 / v;
    if (id32) id32[i] = (float)(1.0 / v);
}

// Right-hand sides. HM_OP_LAMBDA: b = lambda * is_fg (laplacian.py:46-47).
// HM_OP_MASKED: b_U = -(R m) with m = is_fg on known pixels and 0 on known
// rows (estimate_alpha_cf.py:66-74); R's row entries are summed in column
// order like scipy's CSR product.
__global__ void k_rhs_exact(const double* __restrict__ I, const double* __restrict__ wc, int h,
                            int w, const uint8_t* __restrict__ cls, int mode, double lambda,
                            double* __restrict__ b) {
    const long long n = (long long)h * w;
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (mode == HM_OP_LAMBDA) {
        b[i] = lambda * (double)(cls[i] == 1);
        return;
    }
    if (cls[i] != 0) {
        b[i] = 0.0;
        return;
    }
    const int yi = (int)(i / w), xi = (int)(i - (long long)yi * w);
    double sum = 0.0;
    for (int s = 0; s < 25; ++s) {
        const int yj = yi + s / 5 - 2, xj = xi + s % 5 - 2;
        if (yj < 0 || yj >= h || xj < 0 || xj >= w) continue;
        const long long j = (long long)yj * w + xj;
        if (cls[j] == 0) continue;
        sum += entry_exact(I, wc, h, w, yi, xi, yj, xj) * (double)(cls[j] == 1);
    }
    b[i] = -sum;
}

// Value of slot s of row i of the mode-adjusted operator, or 0 outside the image.
__device__ __forceinline__ double slot_value(const double* __restrict__ I,
                                             const double* __restrict__ wc, int h, int w,
                                             const uint8_t* __restrict__ cls, int mode,
                                             double lambda, int yi, int xi, int s, bool* inside,
                                             long long* col) {
    const int yj = yi + s / 5 - 2, xj = xi + s % 5 - 2;
    *inside = yj >= 0 && yj < h && xj >= 0 && xj < w;
    if (!*inside) return 0.0;
    *col = (long long)xj + (long long)yj * w;
    const long long i = (long long)yi * w + xi;
    const bool known_i = cls && cls[i] != 0;
    const bool known_j = cls && cls[*col] != 0;
    if (mode == HM_OP_MASKED && known_i) return s == 12 ? 1.0 : 0.0;
    if (mode == HM_OP_MASKED && known_j) return 0.0;
    double v = entry_exact(I, wc, h, w, yi, xi, yj, xj);
    if (mode == HM_OP_LAMBDA && s == 12 && known_i) v = v + lambda;
    return v;
}

// Canonical CSR (laplacian.py:49-55): in-image slots with a non-zero value,
// in slot order, which is ascending column order.
__global__ void k_system_counts(const double* __restrict__ I, const double* __restrict__ wc,
                                int h, int w, const uint8_t* __restrict__ cls, int mode,
                                double lambda, int64_t* __restrict__ counts) {
    const long long n = (long long)h * w;
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int yi = (int)(i / w), xi = (int)(i - (long long)yi * w);
    int64_t c = 0;
    for (int s = 0; s < 25; ++s) {
        bool inside;
        long long col;
        const double v = slot_value(I, wc, h, w, cls, mode, lambda, yi, xi, s, &inside, &col);
        c += (inside && v != 0.0);
    }
    counts[i] = c;
}

__global__ void k_system_fill(const double* __restrict__ I, const double* __restrict__ wc, int h,
                              int w, const uint8_t* __restrict__ cls, int mode, double lambda,
                              const int64_t* __restrict__ indptr, int64_t* __restrict__ indices,
                              double* __restrict__ values) {
    const long long n = (long long)h * w;
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int yi = (int)(i / w), xi = (int)(i - (long long)yi * w);
    int64_t k = indptr[i];
    for (int s = 0; s < 25; ++s) {
        bool inside;
        long long col;
        const double v = slot_value(I, wc, h, w, cls, mode, lambda, yi, xi, s, &inside, &col);
        if (inside && v != 0.0) {
            indices[k] = col;
            values[k] = v;
            ++k;
        }
    }
}

// ------------------------------------------------------------------ launchers

// ---------------------------------------------------------- any radius r
//
// cf_laplacian.py:5-129 for windows of (2r+1)^2 pixels, as the reference
// assembles them for every radius: per-window data as k_window_exact
// computes it (sums in raster order over the window), and each entry
// L_ij = sum over the windows holding both pixels, in raster order of the
// centre, of delta_ij - (1 + c_i^T M c_j) / (2r+1)^2. This is the general
// (untiled) form; radius 1 runs the tiled kernels above.
__global__ void k_window_exact_r(const double* __restrict__ I, int h, int w, double eps, int r,
                                 const uint8_t* __restrict__ known, double* __restrict__ wc) {
    const long long n = (long long)h * w;
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int y = (int)(t / w), x = (int)(t - (long long)y * w);
    double* flag = wc + 9 * n;
    if (x < r || x > w - 1 - r || y < r || y > h - 1 - r) {
        flag[t] = 0.0;
        return;
    }
    if (known) {  // cf_laplacian.py:34-35
        bool all_known = true;
        for (int dy = -r; dy <= r; ++dy)
            for (int dx = -r; dx <= r; ++dx) all_known &= known[t + (long long)dy * w + dx] != 0;
        if (all_known) {
            flag[t] = 0.0;
            return;
        }
    }
    const int size = 2 * r + 1;
    const double area = (double)(size * size);
    double mean[3];
    for (int c = 0; c < 3; ++c) {
        double sum = 0.0;
        for (int dy = -r; dy <= r; ++dy)
            for (int dx = -r; dx <= r; ++dx) sum += I[c * n + t + (long long)dy * w + dx];
        mean[c] = sum / area;
    }
    double a00 = eps, a01 = 0.0, a02 = 0.0, a11 = eps, a12 = 0.0, a22 = eps;
    for (int dy = -r; dy <= r; ++dy) {
        for (int dx = -r; dx <= r; ++dx) {
            const long long q = t + (long long)dy * w + dx;
            const double c0 = I[q] - mean[0], c1 = I[n + q] - mean[1], c2 = I[2 * n + q] - mean[2];
            a00 += c0 * c0;
            a01 += c0 * c1;
            a02 += c0 * c2;
            a11 += c1 * c1;
            a12 += c1 * c2;
            a22 += c2 * c2;
        }
    }
    a00 /= area; a01 /= area; a02 /= area; a11 /= area; a12 /= area; a22 /= area;
    const double det = a00 * a12 * a12 + a01 * a01 * a22 + a02 * a02 * a11 - a00 * a11 * a22 -
                       2 * a01 * a02 * a12;
    const double inv_det = 1.0 / det;
    wc[t] = mean[0];
    wc[n + t] = mean[1];
    wc[2 * n + t] = mean[2];
    wc[3 * n + t] = (a12 * a12 - a11 * a22) * inv_det;
    wc[4 * n + t] = (a01 * a22 - a02 * a12) * inv_det;
    wc[5 * n + t] = (a02 * a11 - a01 * a12) * inv_det;
    wc[6 * n + t] = (a02 * a02 - a00 * a22) * inv_det;
    wc[7 * n + t] = (a00 * a12 - a01 * a02) * inv_det;
    wc[8 * n + t] = (a01 * a01 - a00 * a11) * inv_det;
    flag[t] = 1.0;
}

// One thread per (pixel i, slot): the reference's CSR layout with
// (4r+1)^2 slots per row, out-of-image slots holding column 0 and value 0.
__global__ void k_assemble_exact_r(const double* __restrict__ I, const double* __restrict__ wc,
                                   int h, int w, int r, double* __restrict__ values,
                                   int64_t* __restrict__ indices, int64_t* __restrict__ indptr) {
    const long long n = (long long)h * w;
    const int span = 4 * r + 1, slots = span * span;
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n * slots) return;
    const long long i = t / slots;
    const int sl = (int)(t - i * slots);
    const int yi = (int)(i / w), xi = (int)(i - (long long)yi * w);
    const int yj = yi + sl / span - 2 * r, xj = xi + sl % span - 2 * r;
    double value = 0.0;
    int64_t col = 0;
    if (yj >= 0 && yj < h && xj >= 0 && xj < w) {
        col = (int64_t)yj * w + xj;
        const double area = (double)((2 * r + 1) * (2 * r + 1));
        const double Ii0 = I[i], Ii1 = I[n + i], Ii2 = I[2 * n + i];
        const double Ij0 = I[col], Ij1 = I[n + col], Ij2 = I[2 * n + col];
        const double delta = (i == col) ? 1.0 : 0.0;
        const int ky0 = max(max(yi, yj) - r, r), ky1 = min(min(yi, yj) + r, h - 1 - r);
        const int kx0 = max(max(xi, xj) - r, r), kx1 = min(min(xi, xj) + r, w - 1 - r);
        for (int ky = ky0; ky <= ky1; ++ky) {
            for (int kx = kx0; kx <= kx1; ++kx) {
                const long long k = (long long)ky * w + kx;
                if (wc[9 * n + k] == 0.0) continue;
                const double mu0 = wc[k], mu1 = wc[n + k], mu2 = wc[2 * n + k];
                const double m00 = wc[3 * n + k], m01 = wc[4 * n + k], m02 = wc[5 * n + k];
                const double m11 = wc[6 * n + k], m12 = wc[7 * n + k], m22 = wc[8 * n + k];
                const double s0 = Ii0 - mu0, s1 = Ii1 - mu1, s2 = Ii2 - mu2;
                const double c0 = m00 * s0 + m01 * s1 + m02 * s2;
                const double c1 = m01 * s0 + m11 * s1 + m12 * s2;
                const double c2 = m02 * s0 + m12 * s1 + m22 * s2;
                const double temp = c0 * (Ij0 - mu0) + c1 * (Ij1 - mu1) + c2 * (Ij2 - mu2);
                value += delta - (1 + temp) / area;
            }
        }
    }
    values[t] = value;
    indices[t] = col;
    if (sl == 0) {
        indptr[i + 1] = (i + 1) * slots;
        if (i == 0) indptr[0] = 0;
    }
}

int cf_exact_setup(hm_cf* op, hipStream_t s) {
    const long long n = op->n;
    const unsigned g = div_up(n, 256);
    k_window_exact<<<g, 256, 0, s>>>(op->I64, op->h, op->w, op->eps, op->cls, op->skip_known,
                                     op->wc);
    HM_LAUNCH_CHECK();
    op->ready = 0;
    return HM_OK;
}

int cf_prepare(hm_cf* op, int need, hipStream_t s) {
    const long long n = op->n;
    if ((need & kCfDiag) && !(op->ready & kCfDiag)) {
        k_diag_exact<<<div_up(n, 256), 256, 0, s>>>(op->I64, op->wc, op->h, op->w, op->cls,
                                                    op->mode, op->lambda, op->diag64, op->diag32,
                                                    op->idiag64, op->idiag32);
        HM_LAUNCH_CHECK();
        op->ready |= kCfDiag;
    }
    if ((need & kCfF32) && !(op->ready & kCfF32)) {
        k_planar_to_f32<<<div_up(9 * n, 256), 256, 0, s>>>(op->wc, op->wc32, 9 * n);
        HM_LAUNCH_CHECK();
        k_planar_to_f32<<<div_up(3 * n, 256), 256, 0, s>>>(op->I64, op->I32, 3 * n);
        HM_LAUNCH_CHECK();
        op->ready |= kCfF32;
    }
    return HM_OK;
}

int cf_upload_image(const double* hwc, hm_cf* op, hipStream_t s) {
    const long long n = op->n;
    k_hwc_to_planar_f64<<<div_up(n, 256), 256, 0, s>>>(hwc, op->I64, n);
    HM_LAUNCH_CHECK();
    op->ready = 0;
    return HM_OK;
}

int cf_assemble(hm_cf* op, double* values, long long stride_i, long long stride_s, int s_lo,
                int64_t* indices, int64_t* indptr, int mode, hipStream_t s) {
    if (s_lo != 0 && (indices || indptr)) {
        set_error("internal: CSR indices need all 25 slots");
        return HM_ERR_ARG;
    }
    k_assemble_exact<<<div_up(op->n, 256), 256, 0, s>>>(op->I64, op->wc, op->h, op->w, op->cls,
                                                         mode, op->lambda, values, stride_i,
                                                         stride_s, s_lo, indices, indptr);
    HM_LAUNCH_CHECK();
    return HM_OK;
}

int cf_assemble_sym(hm_cf* op, hm_dia* A, hipStream_t s) {
    if (A->h != op->h || A->w != op->w) {
        set_error("internal: stored operator shape mismatch");
        return HM_ERR_ARG;
    }
    const dim3 grid(div_up(op->w, kAsmX), div_up(op->h, kAsmY));
    k_assemble_sym<<<grid, kAsmX * kAsmY, 0, s>>>(op->I64, op->wc, op->h, op->w, op->cls,
                                                   op->mode, op->lambda, A->c64, A->c32,
                                                   A->diag64, A->diag32, A->idiag64, A->idiag32);
    HM_LAUNCH_CHECK();
    return HM_OK;
}

int cf_rhs(hm_cf* op, double* b, hipStream_t s) {
    k_rhs_exact<<<div_up(op->n, 256), 256, 0, s>>>(op->I64, op->wc, op->h, op->w, op->cls,
                                                    op->mode, op->lambda, b);
    HM_LAUNCH_CHECK();
    return HM_OK;
}

int trimap_classes(const double* trimap, long long n, double fg, double bg, uint8_t* cls,
                   int64_t* counts, hipStream_t s) {
    const unsigned g = n > 0 ? div_up(n, 256) : 1;
    void* part = nullptr;
    HM_TRY(dev_alloc(&part, sizeof(unsigned long long) * 2 * g));
    k_trimap_classes<<<g, 256, 0, s>>>(trimap, n, fg, bg, cls,
                                       static_cast<unsigned long long*>(part));
    k_trimap_counts<<<1, 256, 0, s>>>(static_cast<unsigned long long*>(part), g, counts);
    const hipError_t e = hipGetLastError();
    dev_free(part);
    if (e != hipSuccess) {
        set_error("trimap classes: %s", hipGetErrorString(e));
        return HM_ERR_HIP;
    }
    return HM_OK;
}

}  // namespace hm

namespace hm {

int cf_system_counts(hm_cf* op, int64_t* counts, hipStream_t s) {
    k_system_counts<<<div_up(op->n, 256), 256, 0, s>>>(op->I64, op->wc, op->h, op->w, op->cls,
                                                        op->mode, op->lambda, counts);
    HM_LAUNCH_CHECK();
    return HM_OK;
}

int cf_system_fill(hm_cf* op, const int64_t* indptr, int64_t* indices, double* values,
                   hipStream_t s) {
    k_system_fill<<<div_up(op->n, 256), 256, 0, s>>>(op->I64, op->wc, op->h, op->w, op->cls,
                                                      op->mode, op->lambda, indptr, indices,
                                                      values);
    HM_LAUNCH_CHECK();
    return HM_OK;
}

}  // namespace hm

namespace {

struct DevTmp {
    void* p = nullptr;
    ~DevTmp() { hm::dev_free(p); }
};

}  // namespace

extern "C" {

int hm_cf_laplacian_csr(const double* image_hwc, int h, int w, double epsilon, int radius,
                        const uint8_t* known, double* values, int64_t* indices, int64_t* indptr,
                        void* stream) {
    using namespace hm;
    if (!image_hwc || !values || !indices || !indptr || h < 1 || w < 1 || radius < 1) {
        set_error("hm_cf_laplacian_csr: invalid argument");
        return HM_ERR_ARG;
    }
    const long long n = (long long)h * w;
    const long long slots = (long long)(4 * radius + 1) * (4 * radius + 1);
    hipStream_t s = static_cast<hipStream_t>(stream);
    DevTmp I, wc;
    HM_TRY(dev_alloc(&I.p, sizeof(double) * 3 * n));
    HM_TRY(dev_alloc(&wc.p, sizeof(double) * 10 * n));
    double* Ip = static_cast<double*>(I.p);
    double* W = static_cast<double*>(wc.p);
    k_hwc_to_planar_f64<<<div_up(n, 256), 256, 0, s>>>(image_hwc, Ip, n);
    HM_LAUNCH_CHECK();
    k_window_exact_r<<<div_up(n, 256), 256, 0, s>>>(Ip, h, w, epsilon, radius, known, W);
    HM_LAUNCH_CHECK();
    k_assemble_exact_r<<<div_up(n * slots, 256), 256, 0, s>>>(Ip, W, h, w, radius, values,
                                                              indices, indptr);
    HM_LAUNCH_CHECK();
    return HM_OK;
}

}  // extern "C"
