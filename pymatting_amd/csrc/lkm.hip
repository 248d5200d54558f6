// Large-kernel matting Laplacian on gfx950 (pymatting/laplacian/lkm_laplacian.py:6-76,
// He et al. 2010): the closed-form Laplacian of (2r+1)^2 windows applied
// matrix-free with box filters, for large radii where the stored form would
// hold (4r+1)^2 entries per row.
//
// Set-up: per window (centres r..h-1-r, r..w-1-r), from box sums of I and
// I I^T: the mean mu, V = (cov + eps/area I)^-1 / area (cov the biased window
// covariance), Vm = V mu and mVm = 1/area + mu . Vm; per pixel, the number
// of windows covering it. Apply (:51-65):
//   P = box_valid(p), Q = box_valid(p I),
//   L p = cnt p + I . box_full(P Vm - V Q) - box_full(P mVm - Q . Vm),
// and the diagonal (:67-74) is box_full(1 - mVm) + I . (2 box_full(Vm) -
// box_full(V) I). Box sums are separable passes that add the 2r+1 taps of an
// output directly (no running sums), so every output is exact to the
// rounding of its own taps. With a class map the operator is
// L + lambda diag(known) (estimate_alpha_lkm.py:51-58).
#include <algorithm>

#include "common.h"
#include "internal.h"

struct hm_lkm {
    int h = 0, w = 0, r = 0;
    long long n = 0, nw = 0;  // pixels, windows
    double lambda = 0.0;
    double* I = nullptr;     // planar [3][n]
    double* win = nullptr;   // [10][nw]: Vm0..2, V00 V01 V02 V11 V12 V22, mVm
    double* cnt = nullptr;   // windows covering each pixel [n]
    double* shift = nullptr; // lambda * known [n], or null
    uint8_t* cls = nullptr;  // owned copy of the class map, or null
    double* tA = nullptr;    // scratch [10][max(Hw*w, h*Ww, n)]
    double* tB = nullptr;    // scratch [10][max(nw, n)]
    long long tA_stride = 0, tB_stride = 0;
};

namespace hm {

namespace {

constexpr int kLkmWin = 10;

// One separable box pass over C channels (planes of H x W, plane stride
// in_stride / out_stride). axis 0 sums along y, axis 1 along x. valid:
// out[t] = sum_{u=t}^{t+2r} in[u] (output side - 2r); full: out[t] =
// sum_{u=t-2r}^{t} in[u] over in-range u (output side + 2r).
__global__ void k_box_pass(const double* __restrict__ in, long long in_stride, int C, int H,
                           int W, int axis, int r, int full, double* __restrict__ out,
                           long long out_stride) {
    const int Ho = axis == 0 ? (full ? H + 2 * r : H - 2 * r) : H;
    const int Wo = axis == 1 ? (full ? W + 2 * r : W - 2 * r) : W;
    const long long m = (long long)Ho * Wo;
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= m * C) return;
    const int c = (int)(t / m);
    const long long o = t - (long long)c * m;
    const int y = (int)(o / Wo), x = (int)(o - (long long)y * Wo);
    const double* src = in + (long long)c * in_stride;
    const int side = axis == 0 ? H : W;
    const int pos = axis == 0 ? y : x;
    const int u0 = full ? pos - 2 * r : pos;
    double s = 0.0;
    for (int u = u0; u <= u0 + 2 * r; ++u) {
        if (u < 0 || u >= side) continue;
        s += axis == 0 ? src[(long long)u * W + x] : src[(long long)y * W + u];
    }
    out[(long long)c * out_stride + o] = s;
}

__global__ void k_hwc_planar(const double* __restrict__ hwc, long long n, double* __restrict__ I) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    I[i] = hwc[3 * i];
    I[n + i] = hwc[3 * i + 1];
    I[2 * n + i] = hwc[3 * i + 2];
}

// q = [I0, I1, I2, I0 I0, I0 I1, I0 I2, I1 I1, I1 I2, I2 I2] (set-up) or
// [p, p I0, p I1, p I2] (apply).
__global__ void k_lkm_products(const double* __restrict__ I, const double* __restrict__ p,
                               long long n, double* __restrict__ q, long long stride) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double a = I[i], b = I[n + i], c = I[2 * n + i];
    if (p) {
        const double v = p[i];
        q[i] = v;
        q[stride + i] = v * a;
        q[2 * stride + i] = v * b;
        q[3 * stride + i] = v * c;
    } else {
        q[i] = a;
        q[stride + i] = b;
        q[2 * stride + i] = c;
        q[3 * stride + i] = a * a;
        q[4 * stride + i] = a * b;
        q[5 * stride + i] = a * c;
        q[6 * stride + i] = b * b;
        q[7 * stride + i] = b * c;
        q[8 * stride + i] = c * c;
    }
}

// Window records from the box sums of I and I I^T (lkm_laplacian.py:38-48).
__global__ void k_lkm_windows(const double* __restrict__ S, long long stride, long long nw,
                              double area, double eps, double* __restrict__ win) {
    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nw) return;
    const double m0 = S[k] / area, m1 = S[stride + k] / area, m2 = S[2 * stride + k] / area;
    const double e = eps / area;
    const double c00 = S[3 * stride + k] / area - m0 * m0 + e;
    const double c01 = S[4 * stride + k] / area - m0 * m1;
    const double c02 = S[5 * stride + k] / area - m0 * m2;
    const double c11 = S[6 * stride + k] / area - m1 * m1 + e;
    const double c12 = S[7 * stride + k] / area - m1 * m2;
    const double c22 = S[8 * stride + k] / area - m2 * m2 + e;
    // Inverse of the symmetric 3x3 matrix by its adjugate.
    const double a00 = c11 * c22 - c12 * c12;
    const double a01 = c02 * c12 - c01 * c22;
    const double a02 = c01 * c12 - c02 * c11;
    const double a11 = c00 * c22 - c02 * c02;
    const double a12 = c01 * c02 - c00 * c12;
    const double a22 = c00 * c11 - c01 * c01;
    const double det = c00 * a00 + c01 * a01 + c02 * a02;
    const double f = 1.0 / (det * area);
    const double v00 = a00 * f, v01 = a01 * f, v02 = a02 * f, v11 = a11 * f, v12 = a12 * f,
                 v22 = a22 * f;
    const double vm0 = v00 * m0 + v01 * m1 + v02 * m2;
    const double vm1 = v01 * m0 + v11 * m1 + v12 * m2;
    const double vm2 = v02 * m0 + v12 * m1 + v22 * m2;
    const double vals[kLkmWin] = {vm0, vm1, vm2, v00, v01, v02, v11, v12, v22,
                                  1.0 / area + (m0 * vm0 + m1 * vm1 + m2 * vm2)};
    for (int q = 0; q < kLkmWin; ++q) win[q * nw + k] = vals[q];
}

// t = [P Vm - V Q (3), P mVm - Q . Vm] per window (lkm_laplacian.py:56-63).
__global__ void k_lkm_window_terms(const double* __restrict__ PQ, long long stride,
                                   const double* __restrict__ win, long long nw,
                                   double* __restrict__ t, long long t_stride) {
    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nw) return;
    const double P = PQ[k], q0 = PQ[stride + k], q1 = PQ[2 * stride + k],
                 q2 = PQ[3 * stride + k];
    const double vm0 = win[k], vm1 = win[nw + k], vm2 = win[2 * nw + k];
    const double v00 = win[3 * nw + k], v01 = win[4 * nw + k], v02 = win[5 * nw + k];
    const double v11 = win[6 * nw + k], v12 = win[7 * nw + k], v22 = win[8 * nw + k];
    const double mvm = win[9 * nw + k];
    t[k] = P * vm0 - (v00 * q0 + v01 * q1 + v02 * q2);
    t[t_stride + k] = P * vm1 - (v01 * q0 + v11 * q1 + v12 * q2);
    t[2 * t_stride + k] = P * vm2 - (v02 * q0 + v12 * q1 + v22 * q2);
    t[3 * t_stride + k] = P * mvm - (q0 * vm0 + q1 * vm1 + q2 * vm2);
}

// y = cnt p + I . T[0..2] - T[3] (+ shift p).
__global__ void k_lkm_combine(const double* __restrict__ I, const double* __restrict__ cnt,
                              const double* __restrict__ shift, const double* __restrict__ p,
                              const double* __restrict__ T, long long stride, long long n,
                              double* __restrict__ y) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double v = cnt[i] * p[i];
    v += I[i] * T[i] + I[n + i] * T[stride + i] + I[2 * n + i] * T[2 * stride + i];
    v -= T[3 * stride + i];
    if (shift) v += shift[i] * p[i];
    y[i] = v;
}

// d = box_full(1 - mVm) + I . (2 box_full(Vm) - box_full(V) I) (+ shift).
__global__ void k_lkm_diag(const double* __restrict__ I, const double* __restrict__ B,
                           long long stride, const double* __restrict__ shift, long long n,
                           double* __restrict__ d) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double a = I[i], b = I[n + i], c = I[2 * n + i];
    const double v00 = B[4 * stride + i], v01 = B[5 * stride + i], v02 = B[6 * stride + i];
    const double v11 = B[7 * stride + i], v12 = B[8 * stride + i], v22 = B[9 * stride + i];
    const double t0 = 2.0 * B[stride + i] - (v00 * a + v01 * b + v02 * c);
    const double t1 = 2.0 * B[2 * stride + i] - (v01 * a + v11 * b + v12 * c);
    const double t2 = 2.0 * B[3 * stride + i] - (v02 * a + v12 * b + v22 * c);
    double v = B[i] + (a * t0 + b * t1 + c * t2);
    if (shift) v += shift[i];
    d[i] = v;
}

// [1 - mVm, Vm0..2, V (6)] per window, the channels the diagonal box-filters.
__global__ void k_lkm_diag_terms(const double* __restrict__ win, long long nw,
                                 double* __restrict__ t, long long t_stride) {
    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nw) return;
    t[k] = 1.0 - win[9 * nw + k];
    for (int q = 0; q < 9; ++q) t[(q + 1) * t_stride + k] = win[q * nw + k];
}

__global__ void k_fill_ones(double* __restrict__ a, long long n) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = 1.0;
}

__global__ void k_lkm_shift(const uint8_t* __restrict__ cls, double lambda, long long n,
                            double* __restrict__ shift) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) shift[i] = cls[i] != 0 ? lambda : 0.0;
}

// b = lambda * is_fg (estimate_alpha_lkm.py:51-53).
__global__ void k_lkm_rhs(const uint8_t* __restrict__ cls, double lambda, long long n,
                          double* __restrict__ b) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = cls[i] == 1 ? lambda : 0.0;
}

__global__ void k_invert(double* __restrict__ d, long long n) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) d[i] = 1.0 / d[i];
}

// Valid (windows from pixels) or full (pixels from windows) 2-D box sum of C
// planes: a vertical pass into tA, then a horizontal pass into dst.
int box2d(const hm_lkm* L, const double* src, long long src_stride, int C, bool full,
          double* dst, long long dst_stride, hipStream_t s) {
    const int r = L->r;
    const int H = full ? L->h - 2 * r : L->h, W = full ? L->w - 2 * r : L->w;
    const int Ho = full ? L->h : L->h - 2 * r;
    const long long m1 = (long long)Ho * W;
    k_box_pass<<<div_up(m1 * C, 256), 256, 0, s>>>(src, src_stride, C, H, W, 0, r, full ? 1 : 0,
                                                   L->tA, L->tA_stride);
    HM_LAUNCH_CHECK();
    const int Wo = full ? L->w : L->w - 2 * r;
    const long long m2 = (long long)Ho * Wo;
    k_box_pass<<<div_up(m2 * C, 256), 256, 0, s>>>(L->tA, L->tA_stride, C, Ho, W, 1, r,
                                                   full ? 1 : 0, dst, dst_stride);
    HM_LAUNCH_CHECK();
    return HM_OK;
}

void release(hm_lkm* L) {
    if (!L) return;
    for (void* p : {(void*)L->I, (void*)L->win, (void*)L->cnt, (void*)L->shift, (void*)L->tA,
                    (void*)L->tB, (void*)L->cls})
        dev_free(p);
    delete L;
}

}  // namespace

int lkm_apply(const hm_lkm* L, const double* x, double* y, hipStream_t s) {
    const long long n = L->n;
    // tB holds [p, p I] over the pixels, then the window sums, then the
    // per-window terms; tA the vertical passes.
    double* q = L->tB;
    k_lkm_products<<<div_up(n, 256), 256, 0, s>>>(L->I, x, n, q, L->tB_stride);
    HM_LAUNCH_CHECK();
    // Window sums go to the upper half of tB's planes (planes 4..7).
    double* PQ = L->tB + 4 * L->tB_stride;
    HM_TRY(box2d(L, q, L->tB_stride, 4, false, PQ, L->tB_stride, s));
    double* terms = L->tB;  // planes 0..3, free again
    k_lkm_window_terms<<<div_up(L->nw, 256), 256, 0, s>>>(PQ, L->tB_stride, L->win, L->nw, terms,
                                                          L->tB_stride);
    HM_LAUNCH_CHECK();
    double* T = L->tB + 4 * L->tB_stride;
    HM_TRY(box2d(L, terms, L->tB_stride, 4, true, T, L->tB_stride, s));
    k_lkm_combine<<<div_up(n, 256), 256, 0, s>>>(L->I, L->cnt, L->shift, x, T, L->tB_stride, n, y);
    HM_LAUNCH_CHECK();
    return HM_OK;
}

long long lkm_rows(const hm_lkm* L) { return L ? L->n : -1; }

}  // namespace hm

extern "C" {

int hm_lkm_create(const double* image_hwc, int h, int w, double epsilon, int radius,
                  const uint8_t* cls, double lambda_value, void* stream, hm_lkm** out) {
    using namespace hm;
    if (!image_hwc || !out || h < 1 || w < 1 || radius < 1) {
        set_error("hm_lkm_create: invalid argument");
        return HM_ERR_ARG;
    }
    if (h < 2 * radius + 1 || w < 2 * radius + 1) {
        // boxfilter.py:140-141 asserts the image holds one window.
        set_error("lkm_laplacian: a %dx%d image is smaller than one %dx%d window", h, w,
                  2 * radius + 1, 2 * radius + 1);
        return HM_ERR_ARG;
    }
    *out = nullptr;
    hipStream_t s = static_cast<hipStream_t>(stream);
    hm_lkm* L = new hm_lkm;
    L->h = h;
    L->w = w;
    L->r = radius;
    L->n = (long long)h * w;
    const int Hw = h - 2 * radius, Ww = w - 2 * radius;
    L->nw = (long long)Hw * Ww;
    L->lambda = lambda_value;
    L->tA_stride = std::max({(long long)Hw * w, (long long)h * Ww, L->n});
    L->tB_stride = std::max(L->nw, L->n);
    const long long n = L->n, nw = L->nw;
    int rc = dev_alloc(reinterpret_cast<void**>(&L->I), sizeof(double) * 3 * n);
    if (rc == HM_OK) rc = dev_alloc(reinterpret_cast<void**>(&L->win), sizeof(double) * 10 * nw);
    if (rc == HM_OK) rc = dev_alloc(reinterpret_cast<void**>(&L->cnt), sizeof(double) * n);
    if (rc == HM_OK) rc = dev_alloc(reinterpret_cast<void**>(&L->tA),
                                    sizeof(double) * 10 * L->tA_stride);
    if (rc == HM_OK) rc = dev_alloc(reinterpret_cast<void**>(&L->tB),
                                    sizeof(double) * 10 * L->tB_stride);
    if (rc == HM_OK && cls) rc = dev_alloc(reinterpret_cast<void**>(&L->shift), sizeof(double) * n);
    if (rc == HM_OK && cls) rc = dev_alloc(reinterpret_cast<void**>(&L->cls), n);
    if (rc != HM_OK) {
        release(L);
        return rc;
    }
    const double area = (double)(2 * radius + 1) * (2 * radius + 1);
    k_hwc_planar<<<div_up(n, 256), 256, 0, s>>>(image_hwc, n, L->I);
    k_lkm_products<<<div_up(n, 256), 256, 0, s>>>(L->I, nullptr, n, L->tB, L->tB_stride);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) {
        // Window sums of the 9 product planes into `win` (stride nw), the
        // records computed from them into tB, then copied back to `win`.
        rc = box2d(L, L->tB, L->tB_stride, 9, false, L->win, nw, s);
        if (rc == HM_OK) {
            k_lkm_windows<<<div_up(nw, 256), 256, 0, s>>>(L->win, nw, nw, area, epsilon, L->tB);
            e = hipGetLastError();
        }
        if (e == hipSuccess && rc == HM_OK)
            e = hipMemcpyAsync(L->win, L->tB, sizeof(double) * 10 * nw, hipMemcpyDeviceToDevice,
                               s);
    }
    // Windows covering each pixel: box_full of ones over the window grid.
    if (e == hipSuccess && rc == HM_OK) {
        k_fill_ones<<<div_up(nw, 256), 256, 0, s>>>(L->tB, nw);
        e = hipGetLastError();
        if (e == hipSuccess) rc = box2d(L, L->tB, nw, 1, true, L->cnt, n, s);
    }
    if (e == hipSuccess && rc == HM_OK && cls) {
        k_lkm_shift<<<div_up(n, 256), 256, 0, s>>>(cls, lambda_value, n, L->shift);
        e = hipGetLastError();
        if (e == hipSuccess) e = hipMemcpyAsync(L->cls, cls, n, hipMemcpyDeviceToDevice, s);
    }
    if (e != hipSuccess) {
        set_error("hm_lkm_create: %s", hipGetErrorString(e));
        rc = HM_ERR_HIP;
    }
    if (rc != HM_OK) {
        release(L);
        return rc;
    }
    *out = L;
    return HM_OK;
}

int hm_lkm_apply(hm_lkm* L, const double* x, double* y, void* stream) {
    if (!L || !x || !y) {
        hm::set_error("hm_lkm_apply: null argument");
        return HM_ERR_ARG;
    }
    return hm::lkm_apply(L, x, y, static_cast<hipStream_t>(stream));
}

int hm_lkm_rhs(hm_lkm* L, double* b, void* stream) {
    using namespace hm;
    if (!L || !b || !L->cls) {
        set_error("hm_lkm_rhs: needs an operator made with a class map");
        return HM_ERR_ARG;
    }
    k_lkm_rhs<<<div_up(L->n, 256), 256, 0, static_cast<hipStream_t>(stream)>>>(L->cls, L->lambda,
                                                                              L->n, b);
    HM_LAUNCH_CHECK();
    return HM_OK;
}

int hm_lkm_diagonal(hm_lkm* L, int inverse, double* d, void* stream) {
    using namespace hm;
    if (!L || !d) {
        set_error("hm_lkm_diagonal: null argument");
        return HM_ERR_ARG;
    }
    hipStream_t s = static_cast<hipStream_t>(stream);
    k_lkm_diag_terms<<<div_up(L->nw, 256), 256, 0, s>>>(L->win, L->nw, L->tB, L->tB_stride);
    HM_LAUNCH_CHECK();
    // The ten full box sums need a pixel-sized plane each (tA holds the
    // vertical pass, tB the window terms): a temporary.
    void* B = nullptr;
    HM_TRY(dev_alloc(&B, sizeof(double) * 10 * L->n));
    int rc = box2d(L, L->tB, L->tB_stride, 10, true, static_cast<double*>(B), L->n, s);
    if (rc == HM_OK) {
        k_lkm_diag<<<div_up(L->n, 256), 256, 0, s>>>(L->I, static_cast<double*>(B), L->n,
                                                     L->shift, L->n, d);
        if (inverse) k_invert<<<div_up(L->n, 256), 256, 0, s>>>(d, L->n);
        if (hipGetLastError() != hipSuccess) {
            set_error("hm_lkm_diagonal: launch failed");
            rc = HM_ERR_HIP;
        }
    }
    dev_free(B);
    return rc;
}

int hm_lkm_destroy(hm_lkm* L) {
    hm::release(L);
    return HM_OK;
}

}  // extern "C"
