/* CPU restatement of pymatting's multilevel foreground estimation.
 *
 * TEST INFRASTRUCTURE / CPU BASELINE ONLY (see oracle/__init__.py).
 *
 * Follows pymatting/foreground/estimate_foreground_ml.py:5-183:
 *   - mean foreground/background colour over alpha > 0.9 / alpha < 0.1,
 *     divided by count + 1e-5 (:78-95);
 *   - a coarse-to-fine pyramid whose level sizes the caller passes in
 *     (round(w0 ** (i / n_levels)), computed in Python, :101-105);
 *   - nearest-neighbour resizes x_src = min(w_src-1, x_dst*w_src/w_dst) (:5-55)
 *     of the full-resolution image/alpha and of the previous level's F/B;
 *   - n_iter in-place raster-order Gauss-Seidel sweeps (:129-175).
 *
 * oracle_fb_ml_f32 mirrors the types Numba infers for the reference: the
 * arrays, a0, a00, da and the b vector are float32, while a1 = 1.0 - a0 and
 * everything derived from it (a01, a11, det, inv_det, F_c) is float64.
 * oracle_fb_ml_f64 runs the same algorithm entirely in double precision.
 *
 * mean_mode 0: float32 running sums in raster order, as the reference;
 * mean_mode 1: double running sums (what the GPU kernels compute).
 * Both divide in double and round the quotient to float32, because Numba
 * selects the float64 ufunc loop for float32_array /= float64_scalar.
 *
 * Build with -ffp-contract=off: the reference performs no FMA contraction.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

static inline int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

#define DEFINE_RESIZE(NAME, T)                                                              \
    static void NAME(T* dst, int hd, int wd, const T* src, int hs, int ws, int depth) {     \
        for (int y = 0; y < hd; ++y) {                                                      \
            int ys = clampi((int)((int64_t)y * hs / hd), 0, hs - 1);                        \
            for (int x = 0; x < wd; ++x) {                                                  \
                int xs = clampi((int)((int64_t)x * ws / wd), 0, ws - 1);                    \
                for (int c = 0; c < depth; ++c)                                             \
                    dst[((int64_t)y * wd + x) * depth + c] =                                \
                        src[((int64_t)ys * ws + xs) * depth + c];                           \
            }                                                                               \
        }                                                                                   \
    }
DEFINE_RESIZE(resize_f32, float)
DEFINE_RESIZE(resize_f64, double)

static const int kDx[4] = {-1, 1, 0, 0};
static const int kDy[4] = {0, 0, -1, 1};

/* One in-place Gauss-Seidel sweep with the reference's float32/float64 mix. */
static void sweep_f32(const float* img, const float* alpha, float* F, float* B, int h, int w,
                      float reg, float gw) {
    for (int y = 0; y < h; ++y) {
        for (int x = 0; x < w; ++x) {
            int64_t p = (int64_t)y * w + x;
            float a0 = alpha[p];
            double a1 = 1.0 - (double)a0;
            float a00 = a0 * a0;
            double a01 = (double)a0 * a1;
            double a11 = a1 * a1;
            float b0[3], b1[3];
            for (int c = 0; c < 3; ++c) {
                b0[c] = a0 * img[p * 3 + c];
                b1[c] = (float)(a1 * (double)img[p * 3 + c]);
            }
            for (int d = 0; d < 4; ++d) {
                int x2 = clampi(x + kDx[d], 0, w - 1);
                int y2 = clampi(y + kDy[d], 0, h - 1);
                int64_t q = (int64_t)y2 * w + x2;
                float gradient = fabsf(a0 - alpha[q]);
                float da = reg + gw * gradient;
                a00 += da;
                a11 += (double)da;
                for (int c = 0; c < 3; ++c) {
                    b0[c] += da * F[q * 3 + c];
                    b1[c] += da * B[q * 3 + c];
                }
            }
            double det = (double)a00 * a11 - a01 * a01;
            double inv_det = 1.0 / det;
            double b00 = inv_det * a11;
            double b01 = inv_det * -a01;
            double b11 = inv_det * (double)a00;
            for (int c = 0; c < 3; ++c) {
                double Fc = b00 * (double)b0[c] + b01 * (double)b1[c];
                double Bc = b01 * (double)b0[c] + b11 * (double)b1[c];
                Fc = Fc < 1.0 ? Fc : 1.0;
                Fc = Fc > 0.0 ? Fc : 0.0;
                Bc = Bc < 1.0 ? Bc : 1.0;
                Bc = Bc > 0.0 ? Bc : 0.0;
                F[p * 3 + c] = (float)Fc;
                B[p * 3 + c] = (float)Bc;
            }
        }
    }
}

static void sweep_f64(const double* img, const double* alpha, double* F, double* B, int h,
                      int w, double reg, double gw) {
    for (int y = 0; y < h; ++y) {
        for (int x = 0; x < w; ++x) {
            int64_t p = (int64_t)y * w + x;
            double a0 = alpha[p];
            double a1 = 1.0 - a0;
            double a00 = a0 * a0;
            double a01 = a0 * a1;
            double a11 = a1 * a1;
            double b0[3], b1[3];
            for (int c = 0; c < 3; ++c) {
                b0[c] = a0 * img[p * 3 + c];
                b1[c] = a1 * img[p * 3 + c];
            }
            for (int d = 0; d < 4; ++d) {
                int x2 = clampi(x + kDx[d], 0, w - 1);
                int y2 = clampi(y + kDy[d], 0, h - 1);
                int64_t q = (int64_t)y2 * w + x2;
                double da = reg + gw * fabs(a0 - alpha[q]);
                a00 += da;
                a11 += da;
                for (int c = 0; c < 3; ++c) {
                    b0[c] += da * F[q * 3 + c];
                    b1[c] += da * B[q * 3 + c];
                }
            }
            double inv_det = 1.0 / (a00 * a11 - a01 * a01);
            double b00 = inv_det * a11, b01 = inv_det * -a01, b11 = inv_det * a00;
            for (int c = 0; c < 3; ++c) {
                double Fc = b00 * b0[c] + b01 * b1[c];
                double Bc = b01 * b0[c] + b11 * b1[c];
                Fc = Fc < 1.0 ? Fc : 1.0;
                Fc = Fc > 0.0 ? Fc : 0.0;
                Bc = Bc < 1.0 ? Bc : 1.0;
                Bc = Bc > 0.0 ? Bc : 0.0;
                F[p * 3 + c] = Fc;
                B[p * 3 + c] = Bc;
            }
        }
    }
}

/* level_h/level_w hold n_levels+1 sizes; F_out/B_out are (h0, w0, 3). */
int oracle_fb_ml_f32(const float* img0, const float* alpha0, int h0, int w0, int n_levels,
                     const int* level_h, const int* level_w, float reg, int n_small, int n_big,
                     int small_size, float gw, int mean_mode, float* F_out, float* B_out) {
    float Fm[3] = {0, 0, 0}, Bm[3] = {0, 0, 0};
    double Fs[3] = {0, 0, 0}, Bs[3] = {0, 0, 0};
    int64_t Fn = 0, Bn = 0;
    for (int64_t p = 0; p < (int64_t)h0 * w0; ++p) {
        if (alpha0[p] > 0.9) {
            for (int c = 0; c < 3; ++c) { Fm[c] += img0[p * 3 + c]; Fs[c] += img0[p * 3 + c]; }
            ++Fn;
        }
        if (alpha0[p] < 0.1) {
            for (int c = 0; c < 3; ++c) { Bm[c] += img0[p * 3 + c]; Bs[c] += img0[p * 3 + c]; }
            ++Bn;
        }
    }
    for (int c = 0; c < 3; ++c) {
        double fsum = mean_mode ? Fs[c] : (double)Fm[c];
        double bsum = mean_mode ? Bs[c] : (double)Bm[c];
        Fm[c] = (float)(fsum / ((double)Fn + 1e-5));
        Bm[c] = (float)(bsum / ((double)Bn + 1e-5));
    }
    int64_t cap = (int64_t)h0 * w0;
    float* img = (float*)malloc(cap * 3 * sizeof(float));
    float* alpha = (float*)malloc(cap * sizeof(float));
    float* F = (float*)malloc(cap * 3 * sizeof(float));
    float* B = (float*)malloc(cap * 3 * sizeof(float));
    float* Fp = (float*)malloc(cap * 3 * sizeof(float));
    float* Bp = (float*)malloc(cap * 3 * sizeof(float));
    int hp = 1, wp = 1;
    memcpy(Fp, Fm, sizeof(Fm));
    memcpy(Bp, Bm, sizeof(Bm));
    for (int l = 0; l <= n_levels; ++l) {
        int h = level_h[l], w = level_w[l];
        resize_f32(img, h, w, img0, h0, w0, 3);
        resize_f32(alpha, h, w, alpha0, h0, w0, 1);
        resize_f32(F, h, w, Fp, hp, wp, 3);
        resize_f32(B, h, w, Bp, hp, wp, 3);
        int n_iter = (w <= small_size && h <= small_size) ? n_small : n_big;
        for (int it = 0; it < n_iter; ++it) sweep_f32(img, alpha, F, B, h, w, reg, gw);
        float* t = Fp; Fp = F; F = t;
        t = Bp; Bp = B; B = t;
        hp = h;
        wp = w;
    }
    memcpy(F_out, Fp, (size_t)hp * wp * 3 * sizeof(float));
    memcpy(B_out, Bp, (size_t)hp * wp * 3 * sizeof(float));
    free(img); free(alpha); free(F); free(B); free(Fp); free(Bp);
    return 0;
}

int oracle_fb_ml_f64(const double* img0, const double* alpha0, int h0, int w0, int n_levels,
                     const int* level_h, const int* level_w, double reg, int n_small, int n_big,
                     int small_size, double gw, double* F_out, double* B_out) {
    double Fm[3] = {0, 0, 0}, Bm[3] = {0, 0, 0};
    int64_t Fn = 0, Bn = 0;
    for (int64_t p = 0; p < (int64_t)h0 * w0; ++p) {
        if (alpha0[p] > 0.9) { for (int c = 0; c < 3; ++c) Fm[c] += img0[p * 3 + c]; ++Fn; }
        if (alpha0[p] < 0.1) { for (int c = 0; c < 3; ++c) Bm[c] += img0[p * 3 + c]; ++Bn; }
    }
    for (int c = 0; c < 3; ++c) {
        Fm[c] /= ((double)Fn + 1e-5);
        Bm[c] /= ((double)Bn + 1e-5);
    }
    int64_t cap = (int64_t)h0 * w0;
    double* img = (double*)malloc(cap * 3 * sizeof(double));
    double* alpha = (double*)malloc(cap * sizeof(double));
    double* F = (double*)malloc(cap * 3 * sizeof(double));
    double* B = (double*)malloc(cap * 3 * sizeof(double));
    double* Fp = (double*)malloc(cap * 3 * sizeof(double));
    double* Bp = (double*)malloc(cap * 3 * sizeof(double));
    int hp = 1, wp = 1;
    memcpy(Fp, Fm, sizeof(Fm));
    memcpy(Bp, Bm, sizeof(Bm));
    for (int l = 0; l <= n_levels; ++l) {
        int h = level_h[l], w = level_w[l];
        resize_f64(img, h, w, img0, h0, w0, 3);
        resize_f64(alpha, h, w, alpha0, h0, w0, 1);
        resize_f64(F, h, w, Fp, hp, wp, 3);
        resize_f64(B, h, w, Bp, hp, wp, 3);
        int n_iter = (w <= small_size && h <= small_size) ? n_small : n_big;
        for (int it = 0; it < n_iter; ++it) sweep_f64(img, alpha, F, B, h, w, reg, gw);
        double* t = Fp; Fp = F; F = t;
        t = Bp; Bp = B; B = t;
        hp = h;
        wp = w;
    }
    memcpy(F_out, Fp, (size_t)hp * wp * 3 * sizeof(double));
    memcpy(B_out, Bp, (size_t)hp * wp * 3 * sizeof(double));
    free(img); free(alpha); free(F); free(B); free(Fp); free(Bp);
    return 0;
}
