// Image gradient of the closed-form operator: the back-propagation step of
// the adjoint method for x = A(I)^-1 b with A = L(I) + lambda diag(known)
// (cf_laplacian.py:5-129 for L; laplacian.py:46-55 for A).
//
// For a loss l(x) with lam = A^-1 dl/dx, dl/dI = -lam^T (dA/dI) x. Only L
// depends on the image. Per window k with pixels i, j:
//   L_k,ij = delta_ij - (1 + c_i^T M c_j) / 9,  c_i = I_i - mu_k,
//   M = (C + eps/9 Id)^-1,  C = (1/9) sum_i c_i c_i^T,
// so lam^T L_k x = (terms free of I) - (1/9) a^T M b with a = sum_i lam_i c_i
// and b = sum_i x_i c_i. Differentiating a, b (through c_i and mu) and M
// (dM = -M dC M, dC = (1/9)(dI_p c_p^T + c_p dI_p^T)) gives, with u = M a,
// v = M b, Lam = sum lam_i, X = sum x_i over the window,
//   dl/dI_p = (1/9) sum_{k containing p} [ (lam_p - Lam/9) v + (x_p - X/9) u
//                                          - (1/9) ((c_p.v) u + (c_p.u) v) ].
#include "internal.h"

namespace hm {

// Per window centre: u = M a, v = M b, Lam/9, X/9 ([8][n]; zero for invalid
// windows, i.e. centres on the border or windows skipped as all-known).
__global__ void k_adj_window(const double* __restrict__ I, const double* __restrict__ wc,
                             const double* __restrict__ lam, const double* __restrict__ x,
                             int h, int w, double* __restrict__ wadj) {
    const long long n = (long long)h * w;
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const double flag = wc[9 * n + t];
    if (flag == 0.0) {
        for (int k = 0; k < 8; ++k) wadj[k * n + t] = 0.0;
        return;
    }
    const double mu0 = wc[t], mu1 = wc[n + t], mu2 = wc[2 * n + t];
    double a0 = 0, a1 = 0, a2 = 0, b0 = 0, b1 = 0, b2 = 0, L = 0, X = 0;
    for (int dy = -1; dy <= 1; ++dy) {
        for (int dx = -1; dx <= 1; ++dx) {
            const long long q = t + (long long)dy * w + dx;
            const double c0 = I[q] - mu0, c1 = I[n + q] - mu1, c2 = I[2 * n + q] - mu2;
            const double lq = lam[q], xq = x[q];
            a0 += lq * c0; a1 += lq * c1; a2 += lq * c2;
            b0 += xq * c0; b1 += xq * c1; b2 += xq * c2;
            L += lq;
            X += xq;
        }
    }
    const double m00 = wc[3 * n + t], m01 = wc[4 * n + t], m02 = wc[5 * n + t];
    const double m11 = wc[6 * n + t], m12 = wc[7 * n + t], m22 = wc[8 * n + t];
    wadj[t] = m00 * a0 + m01 * a1 + m02 * a2;
    wadj[n + t] = m01 * a0 + m11 * a1 + m12 * a2;
    wadj[2 * n + t] = m02 * a0 + m12 * a1 + m22 * a2;
    wadj[3 * n + t] = m00 * b0 + m01 * b1 + m02 * b2;
    wadj[4 * n + t] = m01 * b0 + m11 * b1 + m12 * b2;
    wadj[5 * n + t] = m02 * b0 + m12 * b1 + m22 * b2;
    wadj[6 * n + t] = L / 9.0;
    wadj[7 * n + t] = X / 9.0;
}

// Per pixel: the sum over the (up to nine) windows that contain it, written
// to an HWC float64 gradient (added to it when accumulate != 0).
__global__ void k_adj_pixel(const double* __restrict__ I, const double* __restrict__ wc,
                            const double* __restrict__ wadj, const double* __restrict__ lam,
                            const double* __restrict__ x, int h, int w, int accumulate,
                            double* __restrict__ grad) {
    const long long n = (long long)h * w;
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const int y = (int)(p / w), xx = (int)(p - (long long)y * w);
    const double I0 = I[p], I1 = I[n + p], I2 = I[2 * n + p];
    const double lp = lam[p], xp = x[p];
    double g0 = 0, g1 = 0, g2 = 0;
    for (int dy = -1; dy <= 1; ++dy) {
        const int cy = y + dy;
        if (cy < 1 || cy > h - 2) continue;
        for (int dx = -1; dx <= 1; ++dx) {
            const int cx = xx + dx;
            if (cx < 1 || cx > w - 2) continue;
            const long long k = (long long)cy * w + cx;
            if (wc[9 * n + k] == 0.0) continue;
            const double c0 = I0 - wc[k], c1 = I1 - wc[n + k], c2 = I2 - wc[2 * n + k];
            const double u0 = wadj[k], u1 = wadj[n + k], u2 = wadj[2 * n + k];
            const double v0 = wadj[3 * n + k], v1 = wadj[4 * n + k], v2 = wadj[5 * n + k];
            const double sl = lp - wadj[6 * n + k], sx = xp - wadj[7 * n + k];
            const double cv = (c0 * v0 + c1 * v1 + c2 * v2) / 9.0;
            const double cu = (c0 * u0 + c1 * u1 + c2 * u2) / 9.0;
            g0 += sl * v0 + sx * u0 - (cv * u0 + cu * v0);
            g1 += sl * v1 + sx * u1 - (cv * u1 + cu * v1);
            g2 += sl * v2 + sx * u2 - (cv * u2 + cu * v2);
        }
    }
    double* gp = grad + 3 * p;
    if (accumulate) {
        gp[0] += g0 / 9.0;
        gp[1] += g1 / 9.0;
        gp[2] += g2 / 9.0;
    } else {
        gp[0] = g0 / 9.0;
        gp[1] = g1 / 9.0;
        gp[2] = g2 / 9.0;
    }
}

}  // namespace hm

using namespace hm;

extern "C" int hm_cf_adjoint_image(hm_cf* op, const double* lam, const double* x,
                                   double* grad_hwc, int accumulate, void* stream) {
    if (!op || !lam || !x || !grad_hwc) {
        set_error("hm_cf_adjoint_image: null argument");
        return HM_ERR_ARG;
    }
    hipStream_t s = static_cast<hipStream_t>(stream);
    const long long n = op->n;
    double* wadj = nullptr;
    HM_TRY(dev_alloc(reinterpret_cast<void**>(&wadj), sizeof(double) * 8 * n));
    k_adj_window<<<div_up(n, 256), 256, 0, s>>>(op->I64, op->wc, lam, x, op->h, op->w, wadj);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) {
        k_adj_pixel<<<div_up(n, 256), 256, 0, s>>>(op->I64, op->wc, wadj, lam, x, op->h, op->w,
                                                    accumulate, grad_hwc);
        e = hipGetLastError();
    }
    dev_free(wadj);
    if (e != hipSuccess) {
        set_error("HIP error %s in hm_cf_adjoint_image", hipGetErrorString(e));
        return HM_ERR_HIP;
    }
    return HM_OK;
}
