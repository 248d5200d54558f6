// Tile engine for 25-point grid operators on gfx950.
//
// One 256-thread workgroup owns an output tile of OX x OY grid points. It
// stages the input vector on the tile plus a 2-point halo in LDS (the stencil
// reach of a radius-1 closed-form Laplacian and of its Galerkin coarse
// operators), evaluates A x at every output point, and hands the result to an
// output policy that writes a fused result: A x with <x, A x>, a residual, a
// damped-Jacobi update with <b, z>, or a full-weighting restriction of the
// residual to the next coarser grid.
//
// Operator policies:
//  - CfOp: the matrix-free closed-form Laplacian (Levin et al. 2007) of
//    pymatting/laplacian/cf_laplacian.py:5-129. Per 3x3 window k it recomputes
//    the mean mu_k, the epsilon-regularised covariance and its inverse M_k in
//    LDS, and uses
//       (L x)_i = sum_{k ∋ i} [ x_i - S_k/9 - (I_i - mu_k)^T M_k T_k / 9 ],
//       S_k = sum_{j ∈ k} x_j,  T_k = sum_{j ∈ k} (I_j - mu_k) x_j,
//    which is the row sum of the reference's per-window entries
//    delta_ij - (1 + c_i^T M_k c_j)/9.
//  - DiaOp: a stored 25-point stencil, coefficients slot-major [25][n].
//
// Input policies produce the vector on the halo region: a stored vector, the
// first damped-Jacobi iterate omega*b/D, or a prolongated coarse correction
// added to either.
#pragma once

#include "common.h"

namespace hm {

constexpr int kBlock = 256;

// ----------------------------------------------------------------- operators

struct CfDev {
    int h, w;
    int mode;        // HM_OP_*
    int skip_known;  // drop windows whose 9 pixels are all known
    double eps;
    double lambda;
    const uint8_t* cls;  // nullable: 0 unknown, 1 fg, 2 bg
};

template <typename T, int OX, int OY>
struct CfShared {
    T I[3][OY + 4][OX + 4];
    uint8_t cls[OY + 4][OX + 4];
    T mu[3][OY + 2][OX + 2];
    T v[3][OY + 2][OX + 2];
    T a[OY + 2][OX + 2];
    uint8_t ok[OY + 2][OX + 2];
};

template <typename T>
struct CfOp {
    static constexpr int kProfBase = kProfCfApply;
    CfDev d;
    const T* I;  // planar [3][n]
    int force_mode = -1;  // >= 0 overrides d.mode (the masked right-hand side uses plain L)

    template <int OX, int OY>
    using Shared = CfShared<T, OX, OY>;

    __device__ __forceinline__ int mode() const { return force_mode >= 0 ? force_mode : d.mode; }

    template <int OX, int OY>
    __device__ void load(Shared<OX, OY>& sh, int ox0, int oy0) const {
        constexpr int RX = OX + 4, RY = OY + 4;
        const long long n = (long long)d.h * d.w;
        for (int t = threadIdx.x; t < RX * RY; t += kBlock) {
            const int ly = t / RX, lx = t - ly * RX;
            const int gx = ox0 - 2 + lx, gy = oy0 - 2 + ly;
            const bool in = gx >= 0 && gx < d.w && gy >= 0 && gy < d.h;
            const long long g = (long long)gy * d.w + gx;
#pragma unroll
            for (int c = 0; c < 3; ++c) sh.I[c][ly][lx] = in ? I[c * n + g] : T(0);
            sh.cls[ly][lx] = (in && d.cls) ? d.cls[g] : 0;
        }
    }

    // Window coefficients for all (OX+2) x (OY+2) windows touching the tile.
    template <int OX, int OY>
    __device__ void prepare(Shared<OX, OY>& sh, const T (&sx)[OY + 4][OX + 4], int ox0,
                            int oy0) const {
        constexpr int WX = OX + 2, WY = OY + 2;
        const bool masked = mode() == HM_OP_MASKED;
        for (int t = threadIdx.x; t < WX * WY; t += kBlock) {
            const int wy = t / WX, wx = t - wy * WX;
            const int cx = ox0 - 1 + wx, cy = oy0 - 1 + wy;
            bool valid = cx >= 1 && cx <= d.w - 2 && cy >= 1 && cy <= d.h - 2;
            if (valid && d.skip_known && d.cls) {
                bool all_known = true;
#pragma unroll
                for (int dy = 0; dy < 3; ++dy)
#pragma unroll
                    for (int dx = 0; dx < 3; ++dx) all_known &= sh.cls[wy + dy][wx + dx] != 0;
                valid = !all_known;
            }
            sh.ok[wy][wx] = valid;
            if (!valid) continue;
            T c[3][9];
            T mu[3];
#pragma unroll
            for (int ch = 0; ch < 3; ++ch) {
                T s = 0;
#pragma unroll
                for (int k = 0; k < 9; ++k) s += sh.I[ch][wy + k / 3][wx + k % 3];
                mu[ch] = s / T(9);
#pragma unroll
                for (int k = 0; k < 9; ++k) c[ch][k] = sh.I[ch][wy + k / 3][wx + k % 3] - mu[ch];
            }
            T a00 = T(d.eps), a01 = 0, a02 = 0, a11 = T(d.eps), a12 = 0, a22 = T(d.eps);
#pragma unroll
            for (int k = 0; k < 9; ++k) {
                a00 += c[0][k] * c[0][k];
                a01 += c[0][k] * c[1][k];
                a02 += c[0][k] * c[2][k];
                a11 += c[1][k] * c[1][k];
                a12 += c[1][k] * c[2][k];
                a22 += c[2][k] * c[2][k];
            }
            const T inv9 = T(1) / T(9);
            a00 *= inv9; a01 *= inv9; a02 *= inv9; a11 *= inv9; a12 *= inv9; a22 *= inv9;
            // Inverse by cofactors (cf_laplacian.py:76-92 uses the same closed form).
            const T k00 = a11 * a22 - a12 * a12;
            const T k01 = a02 * a12 - a01 * a22;
            const T k02 = a01 * a12 - a02 * a11;
            const T det = a00 * k00 + a01 * k01 + a02 * k02;
            const T id = T(1) / det;
            const T m00 = k00 * id, m01 = k01 * id, m02 = k02 * id;
            const T m11 = (a00 * a22 - a02 * a02) * id;
            const T m12 = (a01 * a02 - a00 * a12) * id;
            const T m22 = (a00 * a11 - a01 * a01) * id;
            T S = 0, T0 = 0, T1 = 0, T2 = 0;
#pragma unroll
            for (int k = 0; k < 9; ++k) {
                const int ry = wy + k / 3, rx = wx + k % 3;
                T xj = sx[ry][rx];
                if (masked && sh.cls[ry][rx] != 0) xj = 0;
                S += xj;
                T0 += c[0][k] * xj;
                T1 += c[1][k] * xj;
                T2 += c[2][k] * xj;
            }
            sh.mu[0][wy][wx] = mu[0];
            sh.mu[1][wy][wx] = mu[1];
            sh.mu[2][wy][wx] = mu[2];
            sh.v[0][wy][wx] = (m00 * T0 + m01 * T1 + m02 * T2) * inv9;
            sh.v[1][wy][wx] = (m01 * T0 + m11 * T1 + m12 * T2) * inv9;
            sh.v[2][wy][wx] = (m02 * T0 + m12 * T1 + m22 * T2) * inv9;
            sh.a[wy][wx] = S * inv9;
        }
    }

    // (A x) at tile point (lx, ly); xi is the unmasked input at that point.
    template <int OX, int OY>
    __device__ __forceinline__ T apply(const Shared<OX, OY>& sh, const T (&sx)[OY + 4][OX + 4],
                                       int lx, int ly, T xi, long long) const {
        const int m = mode();
        const uint8_t ci = sh.cls[ly + 2][lx + 2];
        if (m == HM_OP_MASKED && ci != 0) return xi;
        const T I0 = sh.I[0][ly + 2][lx + 2], I1 = sh.I[1][ly + 2][lx + 2],
                I2 = sh.I[2][ly + 2][lx + 2];
        T acc = 0;
#pragma unroll
        for (int dy = 0; dy < 3; ++dy) {
#pragma unroll
            for (int dx = 0; dx < 3; ++dx) {
                const int wy = ly + dy, wx = lx + dx;
                if (!sh.ok[wy][wx]) continue;
                acc += xi - sh.a[wy][wx] - (I0 - sh.mu[0][wy][wx]) * sh.v[0][wy][wx] -
                       (I1 - sh.mu[1][wy][wx]) * sh.v[1][wy][wx] -
                       (I2 - sh.mu[2][wy][wx]) * sh.v[2][wy][wx];
            }
        }
        if (m == HM_OP_LAMBDA && ci != 0) acc += T(d.lambda) * xi;
        return acc;
    }
};

template <int OX, int OY>
struct NoShared {};

// Stored 25-point stencil; coef is slot-major [25][n], slot (dy+2)*5 + (dx+2).
template <typename T>
struct DiaOp {
    static constexpr int kProfBase = kProfDiaApply;
    int h, w;
    const T* coef;

    template <int OX, int OY>
    using Shared = NoShared<OX, OY>;

    template <int OX, int OY>
    __device__ void load(Shared<OX, OY>&, int, int) const {}
    template <int OX, int OY>
    __device__ void prepare(Shared<OX, OY>&, const T (&)[OY + 4][OX + 4], int, int) const {}

    template <int OX, int OY>
    __device__ __forceinline__ T apply(const Shared<OX, OY>&, const T (&sx)[OY + 4][OX + 4],
                                       int lx, int ly, T, long long g) const {
        const long long n = (long long)h * w;
        T acc = 0;
#pragma unroll
        for (int s = 0; s < 25; ++s) acc += coef[s * n + g] * sx[ly + s / 5][lx + s % 5];
        return acc;
    }
};

// -------------------------------------------------------------------- inputs

template <typename T>
struct InVec {
    const T* x;
    __device__ __forceinline__ T load(int, int, long long g) const { return x[g]; }
};

template <typename T>
struct InZero {
    __device__ __forceinline__ T load(int, int, long long) const { return T(0); }
};

// First damped-Jacobi iterate from x = 0: omega * b / D (vcycle.py:35-38).
template <typename T>
struct InJac0 {
    const T* b;
    const T* D;
    T omega;
    __device__ __forceinline__ T load(int, int, long long g) const { return omega * b[g] / D[g]; }
};

// base(x) + P e with P = R^T of the full-weighting restriction (vcycle.py:6-29).
template <typename T, class Base>
struct InProlong {
    Base base;
    const T* e;  // coarse vector (h2 x w2)
    int w2, h2;
    __device__ __forceinline__ T load(int gx, int gy, long long g) const {
        T v = base.load(gx, gy, g);
        int jx[2], jy[2];
        T wx[2], wy[2];
        int nx = 0, ny = 0;
        if ((gx & 1) == 0) {
            jx[0] = gx >> 1; wx[0] = T(0.5); nx = 1;
        } else {
            jx[0] = (gx - 1) >> 1; wx[0] = T(0.25);
            jx[1] = (gx + 1) >> 1; wx[1] = T(0.25); nx = 2;
        }
        if ((gy & 1) == 0) {
            jy[0] = gy >> 1; wy[0] = T(0.5); ny = 1;
        } else {
            jy[0] = (gy - 1) >> 1; wy[0] = T(0.25);
            jy[1] = (gy + 1) >> 1; wy[1] = T(0.25); ny = 2;
        }
        T pe = 0;
        for (int a = 0; a < ny; ++a) {
            if (jy[a] < 0 || jy[a] >= h2) continue;
            for (int b = 0; b < nx; ++b) {
                if (jx[b] < 0 || jx[b] >= w2) continue;
                pe += wy[a] * wx[b] * e[(long long)jy[a] * w2 + jx[b]];
            }
        }
        return v + pe;
    }
};

// ------------------------------------------------------------------- outputs

// Reductions write their grid-wide result through this sink.
struct DotSink {
    CgState* st;   // nullable: no reduction
    Reducer red;
    int kind;      // see dot_finalize
};

enum DotKind { kDotPQ = 0, kDotRzFirst = 1, kDotRzUpdate = 2, kDotAux = 3 };

__device__ __forceinline__ void dot_finalize(CgState* st, int kind, double s) {
    if (!isfinite(s)) st->nonfinite = 1;
    switch (kind) {
        case kDotPQ:
            st->pq = s;
            st->alpha = st->rz / s;  // cg.py:90
            break;
        case kDotRzFirst:
            st->rz = s;  // cg.py:80
            break;
        case kDotRzUpdate: {
            double beta = 1.0 / st->rz;  // cg.py:108-110
            st->rz = s;
            st->beta = beta * s;
            break;
        }
        default:
            st->aux = s;
    }
}

// Plain output tiles of OX x OY points starting at (bx*OX, by*OY).
#define HM_TILE_GEOM(OX_, OY_)                                                  \
    static constexpr int OX = OX_, OY = OY_;                                    \
    __device__ __forceinline__ int ox0(int bx) const { return bx * OX; }        \
    __device__ __forceinline__ int oy0(int by) const { return by * OY; }

// y = A x, optionally with <x, A x> (the CG SpMV, cg.py:85-90).
template <typename T>
struct OutApply {
    HM_TILE_GEOM(64, 8)
    static constexpr int kProf = 0;
    T* y;
    DotSink dot;
    __device__ __forceinline__ void point(long long g, int, int, T Ax, T xi, double& acc) const {
        y[g] = Ax;
        acc += (double)xi * (double)Ax;
    }
    __device__ __forceinline__ void outside(int, int) const {}
};

// r = b - A x.
template <typename T>
struct OutResidual {
    HM_TILE_GEOM(64, 8)
    static constexpr int kProf = 3;
    const T* b;
    T* r;
    DotSink dot;
    __device__ __forceinline__ void point(long long g, int, int, T Ax, T, double&) const {
        r[g] = b[g] - Ax;
    }
    __device__ __forceinline__ void outside(int, int) const {}
};

// z = x + omega (b - A x) / D (vcycle.py:42), optionally with <b, z>.
template <typename T>
struct OutSmooth {
    HM_TILE_GEOM(64, 8)
    static constexpr int kProf = 2;
    const T* b;
    const T* D;
    T omega;
    T* z;
    DotSink dot;
    __device__ __forceinline__ void point(long long g, int, int, T Ax, T xi, double& acc) const {
        const T bg = b[g];
        const T zg = xi + omega * (bg - Ax) / D[g];
        z[g] = zg;
        acc += (double)bg * (double)zg;
    }
    __device__ __forceinline__ void outside(int, int) const {}
};

// Residual b - A x on the fine points of a coarse tile, restricted with the
// full-weighting stencil to bc (vcycle.py:77-80).
template <typename T>
struct OutRestrict {
    static constexpr int kProf = 1;
    static constexpr int CX = 32, CY = 4;
    static constexpr int OX = 2 * CX + 1, OY = 2 * CY + 1;
    const T* b;
    T* bc;
    int w2, h2;
    T* sres;  // set by the engine: LDS [OY][OX]
    DotSink dot;
    __device__ __forceinline__ int ox0(int bx) const { return 2 * CX * bx - 1; }
    __device__ __forceinline__ int oy0(int by) const { return 2 * CY * by - 1; }
    __device__ __forceinline__ void point(long long g, int lx, int ly, T Ax, T, double&) const {
        sres[ly * OX + lx] = b[g] - Ax;
    }
    __device__ __forceinline__ void outside(int lx, int ly) const { sres[ly * OX + lx] = T(0); }
    __device__ void restrict_tile(int bx, int by) const {
        for (int t = threadIdx.x; t < CX * CY; t += kBlock) {
            const int cy = t / CX, cx = t - cy * CX;
            const int X = bx * CX + cx, Y = by * CY + cy;
            if (X >= w2 || Y >= h2) continue;
            const int lx = 2 * cx + 1, ly = 2 * cy + 1;
            const T* r0 = sres + (ly - 1) * OX + lx;
            const T* r1 = sres + ly * OX + lx;
            const T* r2 = sres + (ly + 1) * OX + lx;
            T s = (r0[-1] + T(2) * r0[0] + r0[1]) + T(2) * (r1[-1] + T(2) * r1[0] + r1[1]) +
                  (r2[-1] + T(2) * r2[0] + r2[1]);
            bc[(long long)Y * w2 + X] = s * T(0.0625);
        }
    }
};

template <class Out>
struct IsRestrict { static constexpr bool value = false; };
template <typename T>
struct IsRestrict<OutRestrict<T>> { static constexpr bool value = true; };

// -------------------------------------------------------------------- kernel

template <typename T, class Op, class In, class Out>
__global__ __launch_bounds__(kBlock) void tile_kernel(Op op, In in, Out out, int h, int w,
                                                      const CgState* guard) {
    if (solve_done(guard)) return;
    constexpr int OX = Out::OX, OY = Out::OY;
    constexpr int RX = OX + 4, RY = OY + 4;
    __shared__ T sx[RY][RX];
    __shared__ typename Op::template Shared<OX, OY> sh;
    __shared__ T sres[IsRestrict<Out>::value ? OX * OY : 1];
    const int ox0 = out.ox0(blockIdx.x), oy0 = out.oy0(blockIdx.y);

    for (int t = threadIdx.x; t < RX * RY; t += kBlock) {
        const int ly = t / RX, lx = t - ly * RX;
        const int gx = ox0 - 2 + lx, gy = oy0 - 2 + ly;
        const bool inb = gx >= 0 && gx < w && gy >= 0 && gy < h;
        sx[ly][lx] = inb ? in.load(gx, gy, (long long)gy * w + gx) : T(0);
    }
    op.template load<OX, OY>(sh, ox0, oy0);
    __syncthreads();
    op.template prepare<OX, OY>(sh, sx, ox0, oy0);
    __syncthreads();

    Out o = out;
    if constexpr (IsRestrict<Out>::value) o.sres = sres;
    double acc = 0.0;
    for (int t = threadIdx.x; t < OX * OY; t += kBlock) {
        const int ly = t / OX, lx = t - ly * OX;
        const int gx = ox0 + lx, gy = oy0 + ly;
        if (gx < 0 || gx >= w || gy < 0 || gy >= h) {
            o.outside(lx, ly);
            continue;
        }
        const long long g = (long long)gy * w + gx;
        const T xi = sx[ly + 2][lx + 2];
        const T Ax = op.template apply<OX, OY>(sh, sx, lx, ly, xi, g);
        o.point(g, lx, ly, Ax, xi, acc);
    }
    if constexpr (IsRestrict<Out>::value) {
        __syncthreads();
        o.restrict_tile(blockIdx.x, blockIdx.y);
    }
    if (o.dot.st != nullptr) {
        double v[1] = {acc}, tot[1];
        if (grid_sum<kBlock, 1>(v, o.dot.red, tot)) dot_finalize(o.dot.st, o.dot.kind, tot[0]);
    }
}

template <class Out>
inline dim3 tile_grid(const Out&, int h, int w, int h2, int w2) {
    if constexpr (IsRestrict<Out>::value)
        return dim3(div_up(w2, Out::CX), div_up(h2, Out::CY));
    else
        return dim3(div_up(w, Out::OX), div_up(h, Out::OY));
}

}  // namespace hm
