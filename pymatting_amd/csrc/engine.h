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
//    pymatting/laplacian/cf_laplacian.py:5-129. The mean mu_k and inverse
//    regularised covariance M_k of every 3x3 window are computed once per
//    operator (in the reference's operation order; float32 operators round
//    the float64 values); each pass stages the image and vector in LDS and
//    uses
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

// LDS records: per region pixel {I0, I1, I2, x (masked)}, per window
// {S/9, mu0, mu1, mu2, v0, v1, v2, valid}; read as 16-byte vectors.
template <typename T>
struct alignas(4 * sizeof(T)) Vec4 {
    T a, b, c, d;
};

template <typename T, int OX, int OY>
struct CfShared {
    Vec4<T> px[OY + 4][OX + 4];
    Vec4<T> win0[OY + 2][OX + 2];  // {S/9, mu0, mu1, mu2}
    Vec4<T> win1[OY + 2][OX + 2];  // {v0, v1, v2, valid}
    uint8_t cls[OY + 4][OX + 4];
};

// Window coefficients held in registers from the start of the kernel.
// Both precisions store the window mean and inverse covariance (9 values),
// float32 ones rounded from the float64 values: recomputing the mean in
// float32 perturbs nearly singular windows far more than the rounding.
// Rec: the windows are recomputed from the staged image instead (below), and
// only their geometry is kept.
template <typename T, int OX, int OY, bool Rec = false>
struct CfPrefetch {
    static constexpr int NW = ((OX + 2) * (OY + 2) + kBlock - 1) / kBlock;
    static constexpr int NQ = Rec ? 0 : 9;
    T cf[NW][Rec ? 1 : 9];
    bool geo[NW];
};

// Mean and inverse regularised covariance of a 3x3 window from its nine
// (float32) pixels, in float64 and in the operation order of
// cf_laplacian.py:38-92 (the order of the stored window planes): window sum
// and mean, centred colours, covariance with eps, the negated determinant and
// its cofactors, no fused multiply-adds. Windows with two eigenvalues near
// eps / 9 make the inverse sensitive to the last bit of the float64
// arithmetic (a product with 1/9 in place of the division moves entries by
// 1e-5 relative there), so any other order would give a different operator
// from the stored and assembled ones; with this one, an image whose values
// are float32 numbers gets exactly the stored windows.
template <typename T>
__device__ __forceinline__ void cf_window_f64(const Vec4<T> (&p)[9], double eps, T (&mu)[3],
                                              T (&m)[6]) {
#pragma clang fp contract(off)
    double s0 = 0.0, s1 = 0.0, s2 = 0.0;
#pragma unroll
    for (int j = 0; j < 9; ++j) {
        s0 += (double)p[j].a;
        s1 += (double)p[j].b;
        s2 += (double)p[j].c;
    }
    const double u0 = s0 / 9, u1 = s1 / 9, u2 = s2 / 9;
    double a00 = eps, a01 = 0.0, a02 = 0.0, a11 = eps, a12 = 0.0, a22 = eps;
#pragma unroll
    for (int j = 0; j < 9; ++j) {
        const double c0 = (double)p[j].a - u0, c1 = (double)p[j].b - u1;
        const double c2 = (double)p[j].c - u2;
        a00 += c0 * c0;
        a01 += c0 * c1;
        a02 += c0 * c2;
        a11 += c1 * c1;
        a12 += c1 * c2;
        a22 += c2 * c2;
    }
    a00 /= 9;
    a01 /= 9;
    a02 /= 9;
    a11 /= 9;
    a12 /= 9;
    a22 /= 9;
    const double det = a00 * a12 * a12 + a01 * a01 * a22 + a02 * a02 * a11 - a00 * a11 * a22 -
                       2 * a01 * a02 * a12;
    const double inv_det = 1.0 / det;
    mu[0] = (T)u0;
    mu[1] = (T)u1;
    mu[2] = (T)u2;
    m[0] = (T)((a12 * a12 - a11 * a22) * inv_det);
    m[1] = (T)((a01 * a22 - a02 * a12) * inv_det);
    m[2] = (T)((a02 * a11 - a01 * a12) * inv_det);
    m[3] = (T)((a02 * a02 - a00 * a22) * inv_det);
    m[4] = (T)((a00 * a12 - a01 * a02) * inv_det);
    m[5] = (T)((a01 * a01 - a00 * a11) * inv_det);
}

// a * b + c with one rounding, written out so that every instantiation of
// the matrix-free operator rounds alike (the compiler's own contraction
// choices depend on where the operands come from, and the operator's nearly
// singular windows turn a different choice into a visible difference).
__device__ __forceinline__ float fmadd(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double fmadd(double a, double b, double c) {
    return __builtin_fma(a, b, c);
}

// ------------------------------------------------ V-cycle pointwise arithmetic
//
// The V-cycle's per-point formulas (vcycle.py:32-43 and 77-100), each written
// out operation by operation with no implicit contraction. Every kernel that
// evaluates them -- the one-level tile kernels and the two-level kernels that
// recompute an intermediate level on a halo -- therefore rounds alike, and
// the two launch structures give bit-identical V-cycles.

// First damped-Jacobi iterate from zero: omega * b / D, with iD = 1 / D.
template <typename T>
__device__ __forceinline__ T vc_jac0(T omega, T b, T iD) {
#pragma clang fp contract(off)
    return omega * b * iD;
}

// v + (P e) at a fine point of parities (ox, oy): P e takes 1/2 of the
// coincident coarse value per axis at even coordinates and 1/4 of each of the
// two neighbours at odd ones. e00 is the coarse value at (x >> 1, y >> 1),
// e01 the one to its right, e10 below, e11 diagonal (zero outside the coarse
// grid; taps that a parity does not use are ignored).
template <typename T>
__device__ __forceinline__ T vc_prolong(T v, bool ox, bool oy, T e00, T e01, T e10, T e11) {
#pragma clang fp contract(off)
    const T wx0 = ox ? T(0.25) : T(0.5), wy0 = oy ? T(0.25) : T(0.5);
    T row0 = wx0 * e00;
    if (ox) row0 = fmadd(T(0.25), e01, row0);
    T pe = wy0 * row0;
    if (oy) {
        T row1 = wx0 * e10;
        if (ox) row1 = fmadd(T(0.25), e11, row1);
        pe = fmadd(T(0.25), row1, pe);
    }
    return v + pe;
}

// One damped-Jacobi sweep at a point: y + omega (b - A y) / D.
template <typename T>
__device__ __forceinline__ T vc_smooth(T y, T b, T Ay, T iD, T omega) {
#pragma clang fp contract(off)
    return fmadd(omega * (b - Ay), iD, y);
}

// Full-weighting restriction of a 3 x 3 block of fine residuals; r0, r1, r2
// point at the centre column of its three rows.
template <typename T>
__device__ __forceinline__ T vc_restrict9(const T* r0, const T* r1, const T* r2) {
#pragma clang fp contract(off)
    const T a = fmadd(T(2), r0[0], r0[-1]) + r0[1];
    const T b = fmadd(T(2), r1[0], r1[-1]) + r1[1];
    const T c = fmadd(T(2), r2[0], r2[-1]) + r2[1];
    return (fmadd(T(2), b, a) + c) * T(0.0625);
}

// 25-tap stencil sum sum_s c[s] x[s] as a chain of fused multiply-adds in
// slot order; x(s) returns the input under slot s.
template <typename T, class C, class X>
__device__ __forceinline__ T vc_stencil(const C& c, const X& x) {
    T acc = 0;
#pragma unroll
    for (int s = 0; s < 25; ++s) acc = fmadd(c(s), x(s), acc);
    return acc;
}

// Rec = false: the window data are read from W. Rec = true (float32 passes):
// every pass recomputes each window's mean and inverse covariance from the
// image it stages anyway (cf_window_f64), so a pass reads the image, the
// class map and its vectors and no window planes.
template <typename T, bool Rec = false>
struct CfOp {
    static constexpr int kProfBase = Rec ? kProfCfRecApply : kProfCfApply;
    static constexpr bool kPointPrefetch = false;
    CfDev d;
    const T* I;  // planar [3][n]
    // Per-window data at the centre pixel [9][n]: mean (3), inverse covariance (6).
    const T* W;
    int force_mode = -1;  // >= 0 overrides d.mode (the masked right-hand side uses plain L)

    template <int OX, int OY>
    using Shared = CfShared<T, OX, OY>;

    __device__ __forceinline__ int mode() const { return force_mode >= 0 ? force_mode : d.mode; }

    // Stages the image, the (masked) input and the class map at region point
    // (ly, lx). In masked mode known pixels enter the window sums as 0.
    template <int OX, int OY>
    __device__ __forceinline__ void load_point(Shared<OX, OY>& sh, int ly, int lx, bool in,
                                               long long g, T x) const {
        const long long n = (long long)d.h * d.w;
        const uint8_t c = (in && d.cls) ? d.cls[g] : 0;
        Vec4<T> v;
        v.a = in ? I[g] : T(0);
        v.b = in ? I[n + g] : T(0);
        v.c = in ? I[2 * n + g] : T(0);
        v.d = (mode() == HM_OP_MASKED && c != 0) ? T(0) : x;
        sh.px[ly][lx] = v;
        sh.cls[ly][lx] = c;
    }

    template <int OX, int OY>
    using Prefetch = CfPrefetch<T, OX, OY, Rec>;

    // Loads the coefficients of this thread's windows into registers; issued
    // before the region loads so both latencies overlap.
    template <int OX, int OY>
    __device__ __forceinline__ void prefetch(Prefetch<OX, OY>& pf, int ox0, int oy0) const {
        constexpr int WX = OX + 2, WY = OY + 2;
        constexpr int NW = Prefetch<OX, OY>::NW, NQ = Prefetch<OX, OY>::NQ;
        const long long n = (long long)d.h * d.w;
#pragma unroll
        for (int k = 0; k < NW; ++k) {
            const int t = threadIdx.x + k * kBlock;
            const int wy = t / WX, wx = t - wy * WX;
            const int cx = ox0 - 1 + wx, cy = oy0 - 1 + wy;
            const bool ok = t < WX * WY && cx >= 1 && cx <= d.w - 2 && cy >= 1 && cy <= d.h - 2;
            pf.geo[k] = ok;
            const long long gc = ok ? (long long)cy * d.w + cx : 0;
            (void)gc;
            (void)n;
#pragma unroll
            for (int q = 0; q < NQ; ++q) pf.cf[k][q] = ok ? W[q * n + gc] : T(0);
        }
    }

    // Window records for all (OX+2) x (OY+2) windows touching the tile.
    template <int OX, int OY>
    __device__ void prepare(Shared<OX, OY>& sh, const Prefetch<OX, OY>& pf) const {
#pragma clang fp contract(off)
        constexpr int WX = OX + 2, WY = OY + 2;
        constexpr int NW = Prefetch<OX, OY>::NW, NQ = Prefetch<OX, OY>::NQ;
        constexpr int M0 = NQ - 6;  // index of m00
#pragma unroll
        for (int k = 0; k < NW; ++k) {
            const int t = threadIdx.x + k * kBlock;
            if (t >= WX * WY) continue;
            const int wy = t / WX, wx = t - wy * WX;
            bool valid = pf.geo[k];
            if (valid && d.skip_known && d.cls) {
                bool all_known = true;
#pragma unroll
                for (int dy = 0; dy < 3; ++dy)
#pragma unroll
                    for (int dx = 0; dx < 3; ++dx) all_known &= sh.cls[wy + dy][wx + dx] != 0;
                valid = !all_known;
            }
            Vec4<T> p[9];
#pragma unroll
            for (int j = 0; j < 9; ++j) p[j] = sh.px[wy + j / 3][wx + j % 3];
            T mu[3], mw[6];
            if constexpr (Rec) {
                cf_window_f64<T>(p, d.eps, mu, mw);
            } else if constexpr (NQ == 9) {
                mu[0] = pf.cf[k][0];
                mu[1] = pf.cf[k][1];
                mu[2] = pf.cf[k][2];
            } else {
                const T inv9 = T(1) / T(9);
                T s0 = 0, s1 = 0, s2 = 0;
#pragma unroll
                for (int j = 0; j < 9; ++j) {
                    s0 += p[j].a;
                    s1 += p[j].b;
                    s2 += p[j].c;
                }
                mu[0] = s0 * inv9;
                mu[1] = s1 * inv9;
                mu[2] = s2 * inv9;
            }
            const T inv9 = T(1) / T(9);
            T S = 0, T0 = 0, T1 = 0, T2 = 0;
#pragma unroll
            for (int j = 0; j < 9; ++j) {
                S += p[j].d;
                T0 = fmadd(p[j].a - mu[0], p[j].d, T0);
                T1 = fmadd(p[j].b - mu[1], p[j].d, T1);
                T2 = fmadd(p[j].c - mu[2], p[j].d, T2);
            }
            // Invalid windows get all-zero records, so the output loop needs
            // no branch.
            const T okf = valid ? T(1) : T(0);
            if constexpr (!Rec) {
#pragma unroll
                for (int q = 0; q < 6; ++q) mw[q] = pf.cf[k][M0 + q];
            }
            const T m00 = mw[0], m01 = mw[1], m02 = mw[2];
            const T m11 = mw[3], m12 = mw[4], m22 = mw[5];
            Vec4<T> r0, r1;
            r0.a = okf * S * inv9;
            r0.b = mu[0];
            r0.c = mu[1];
            r0.d = mu[2];
            r1.a = okf * fmadd(m02, T2, fmadd(m01, T1, m00 * T0)) * inv9;
            r1.b = okf * fmadd(m12, T2, fmadd(m11, T1, m01 * T0)) * inv9;
            r1.c = okf * fmadd(m22, T2, fmadd(m12, T1, m02 * T0)) * inv9;
            r1.d = okf;
            sh.win0[wy][wx] = r0;
            sh.win1[wy][wx] = r1;
        }
    }

    // (A x) at tile point (lx, ly); xi is the unmasked input at that point.
    template <int OX, int OY>
    __device__ __forceinline__ T apply(const Shared<OX, OY>& sh, const Prefetch<OX, OY>&,
                                       const T (&)[OY + 4][OX + 4], int, int lx, int ly, T xi,
                                       long long) const {
#pragma clang fp contract(off)
        const int m = mode();
        const uint8_t ci = sh.cls[ly + 2][lx + 2];
        const Vec4<T> p = sh.px[ly + 2][lx + 2];
        T acc = 0, cnt = 0;
#pragma unroll
        for (int dy = 0; dy < 3; ++dy) {
#pragma unroll
            for (int dx = 0; dx < 3; ++dx) {
                const Vec4<T> r0 = sh.win0[ly + dy][lx + dx];
                const Vec4<T> r1 = sh.win1[ly + dy][lx + dx];
                acc -= fmadd(p.c - r0.d, r1.c,
                             fmadd(p.b - r0.c, r1.b, fmadd(p.a - r0.b, r1.a, r0.a)));
                cnt += r1.d;
            }
        }
        // Invalid windows have zero records; each valid one adds x_i.
        acc = fmadd(cnt, xi, acc);
        if (m == HM_OP_MASKED && ci != 0) return xi;
        if (m == HM_OP_LAMBDA && ci != 0) acc = fmadd(T(d.lambda), xi, acc);
        return acc;
    }
};

// Stored symmetric 25-point stencils keep the 13 slots s = (dy+2)*5 + (dx+2)
// with s >= 12 ((dy, dx) >= (0, 0) in raster order), slot-major [13][n]; the
// coefficient of slot s < 12 at point i is that of slot 24 - s at the
// neighbour i + (dy, dx) (A is symmetric). The neighbour index is clamped into
// the array: callers multiply out-of-grid couplings by zero.
constexpr int kSymSlots = 13;

template <typename T>
__device__ __forceinline__ T sym_coef(const T* __restrict__ c, long long n, int w, long long i,
                                      int s) {
    if (s >= 12) return c[(s - 12) * n + i];
    long long j = i + (long long)(s / 5 - 2) * w + (s % 5 - 2);
    j = j < 0 ? 0 : (j >= n ? n - 1 : j);
    return c[(12 - s) * n + j];
}

// sym_coef for grids below 2^31 points with 32-bit indices: the plane base
// is uniform across the workgroup and the neighbour index only needs the
// lower clamp (slots s < 12 point up or left).
template <typename T>
__device__ __forceinline__ T sym_coef32(const T* __restrict__ c, int n, int w, int i, int s) {
    if (s >= 12) return (c + (size_t)(s - 12) * n)[(unsigned)i];
    const int j = i + (s / 5 - 2) * w + (s % 5 - 2);
    return (c + (size_t)(12 - s) * n)[(unsigned)(j < 0 ? 0 : j)];
}

template <int OX, int OY>
struct NoShared {};

// Coefficients of this thread's output points, loaded before the tile
// barrier so their latency overlaps the halo loads (PF = true: small grids,
// whose launches are latency-bound). PF = false loads them at use, which keeps
// register pressure and occupancy where the large grids need them.
template <typename T, int OX, int OY, bool PF>
struct DiaPrefetch {
    static constexpr int NO = (OX * OY + kBlock - 1) / kBlock;
    T c[NO][25];
};
template <typename T, int OX, int OY>
struct DiaPrefetch<T, OX, OY, false> {};

template <typename T, bool PF = false>
struct DiaOp {
    static constexpr int kProfBase = kProfDiaApply;
    static constexpr bool kPointPrefetch = PF;
    int h, w;
    const T* coef;  // [13][n] symmetric storage

    template <int OX, int OY>
    using Shared = NoShared<OX, OY>;
    template <int OX, int OY>
    using Prefetch = DiaPrefetch<T, OX, OY, PF>;

    template <int OX, int OY>
    __device__ __forceinline__ void prefetch(Prefetch<OX, OY>& pf, int ox0, int oy0) const {
        if constexpr (PF) {
            const int n = h * w;
            constexpr int NO = (OX * OY + kBlock - 1) / kBlock;
#pragma unroll
            for (int k = 0; k < NO; ++k) {
                const int t = threadIdx.x + k * kBlock;
                const int ly = t / OX, lx = t - ly * OX;
                const int gx = ox0 + lx, gy = oy0 + ly;
                const bool in = t < OX * OY && gx >= 0 && gx < w && gy >= 0 && gy < h;
                const int g = in ? gy * w + gx : 0;
#pragma unroll
                for (int s = 0; s < 25; ++s) pf.c[k][s] = in ? sym_coef32(coef, n, w, g, s) : T(0);
            }
        }
    }
    template <int OX, int OY>
    __device__ __forceinline__ void load_point(Shared<OX, OY>&, int, int, bool, long long, T) const {}
    template <int OX, int OY>
    __device__ void prepare(Shared<OX, OY>&, const Prefetch<OX, OY>&) const {}

    template <int OX, int OY>
    __device__ __forceinline__ T apply(const Shared<OX, OY>&, const Prefetch<OX, OY>& pf,
                                       const T (&sx)[OY + 4][OX + 4], int k, int lx, int ly, T,
                                       long long g) const {
        const auto xs = [&](int s) { return sx[ly + s / 5][lx + s % 5]; };
        if constexpr (PF) {
            return vc_stencil<T>([&](int s) { return pf.c[k][s]; }, xs);
        } else {
            const int n = h * w;
            const int i = (int)g;
            return vc_stencil<T>([&](int s) { return sym_coef32(coef, n, w, i, s); }, xs);
        }
    }
};

// -------------------------------------------------------------------- inputs

template <typename T>
struct InVec {
    const T* x;
    __device__ __forceinline__ T load(int, int, long long g) const { return x[g]; }
};

// p = z + beta p (cg.py:112-113) formed on the fly from the previous
// direction; the apply kernel writes the tile's new p to another buffer.
template <typename T>
struct InPupdate {
    const T* z;
    const T* p_old;
    const CgState* st;
    __device__ __forceinline__ T load(int, int, long long g) const {
        return p_old[g] * (T)st->beta + z[g];
    }
};

template <typename T>
struct InZero {
    __device__ __forceinline__ T load(int, int, long long) const { return T(0); }
};

// First damped-Jacobi iterate from x = 0: omega * b / D (vcycle.py:35-38).
template <typename T>
struct InJac0 {
    const T* b;
    const T* iD;  // 1 / D
    T omega;
    __device__ __forceinline__ T load(int, int, long long g) const {
        return vc_jac0(omega, b[g], iD[g]);
    }
};

// Inputs that stage data in LDS before the tile's loads (InProlong) provide
// Shared<RX, RY>, stage() and load_staged(); the others are read directly.
template <class In>
struct InStaged { static constexpr bool value = false; };

// base(x) + P e with P = R^T of the full-weighting restriction (vcycle.py:6-29).
template <typename T, class Base>
struct InProlong {
    Base base;
    const T* e;  // coarse vector (h2 x w2)
    int w2, h2;
    // The coarse block under a tile's input region of RX x RY fine points
    // starting at the even (ox0 - 2, oy0 - 2), with zeros outside the coarse
    // grid; each coarse value is read from memory once per tile.
    template <int RX, int RY>
    struct Shared {
        static constexpr int SW = RX / 2 + 2, SH = RY / 2 + 2;
        T e[SH][SW];
    };
    template <int RX, int RY>
    __device__ __forceinline__ void stage(Shared<RX, RY>& sh, int ox0, int oy0) const {
        using S = Shared<RX, RY>;
        const int cx0 = (ox0 - 2) >> 1, cy0 = (oy0 - 2) >> 1;
        for (int t = threadIdx.x; t < S::SW * S::SH; t += kBlock) {
            const int ly = t / S::SW, lx = t - ly * S::SW;
            const int cx = cx0 + lx, cy = cy0 + ly;
            const bool v = cx >= 0 && cx < w2 && cy >= 0 && cy < h2;
            sh.e[ly][lx] = v ? e[(long long)cy * w2 + cx] : T(0);
        }
    }
    template <int RX, int RY>
    __device__ __forceinline__ T load_staged(const Shared<RX, RY>& sh, int ox0, int oy0, int gx,
                                             int gy, long long g) const {
        const T v = base.load(gx, gy, g);
        // Same taps and weights as load(); zeros stand for the dropped taps.
        const int lx0 = (gx >> 1) - ((ox0 - 2) >> 1), ly0 = (gy >> 1) - ((oy0 - 2) >> 1);
        const bool ox = gx & 1, oy = gy & 1;
        return vc_prolong(v, ox, oy, sh.e[ly0][lx0], ox ? sh.e[ly0][lx0 + 1] : T(0),
                          oy ? sh.e[ly0 + 1][lx0] : T(0),
                          (ox && oy) ? sh.e[ly0 + 1][lx0 + 1] : T(0));
    }
    __device__ __forceinline__ T load(int gx, int gy, long long g) const {
        const T v = base.load(gx, gy, g);
        // Fine (gx, gy) takes 1/2 of the coincident coarse point per axis when
        // even, 1/4 of each of the two neighbours when odd (x0 = gx >> 1
        // floors, so the odd case's taps are x0 and x0 + 1); taps outside the
        // coarse grid are dropped (read as zero).
        const int x0 = gx >> 1, y0 = gy >> 1;
        const bool ox = gx & 1, oy = gy & 1;
        const bool vx0 = x0 >= 0 && x0 < w2, vx1 = ox && x0 + 1 < w2;
        const bool vy0 = y0 >= 0 && y0 < h2, vy1 = oy && y0 + 1 < h2;
        const long long r0 = (long long)y0 * w2, r1 = r0 + w2;
        const T e00 = (vy0 && vx0) ? e[r0 + x0] : T(0);
        const T e01 = (vy0 && vx1) ? e[r0 + x0 + 1] : T(0);
        const T e10 = (vy1 && vx0) ? e[r1 + x0] : T(0);
        const T e11 = (vy1 && vx1) ? e[r1 + x0 + 1] : T(0);
        return vc_prolong(v, ox, oy, e00, e01, e10, e11);
    }
};

template <typename T, class Base>
struct InStaged<InProlong<T, Base>> { static constexpr bool value = true; };

template <class In, int RX, int RY, bool = InStaged<In>::value>
struct InShared { struct type {}; };
template <class In, int RX, int RY>
struct InShared<In, RX, RY, true> { using type = typename In::template Shared<RX, RY>; };

// ------------------------------------------------------------------- outputs

// Reductions write their grid-wide result through this sink.
struct DotSink {
    CgState* st;   // nullable: no reduction
    Reducer red;
    int kind;      // see dot_finalize
    bool defer = false;  // partials only: the consuming kernel sums them
};

enum DotKind {
    kDotPQ = 0, kDotRzFirst = 1, kDotRzUpdate = 2, kDotAux = 3, kDotUpdate = 4, kDotReorth = 5,
    // <r, z> whose finalisation also completes the preceding CG update (its
    // ||r||^2 partials in slot 2 of the same reducer): one finalize launch
    // per iteration instead of two.
    kDotUpdateRz = 6,
    // Start of a solve without a host round trip (cg.py:66-76): ||b|| of a
    // solve from x = 0 (r = b; b = 0 is solved by x = 0 and freezes the
    // solve), ||b|| alone, and the initial residual test.
    kDotNormB = 7, kDotNormBOnly = 8, kDotInitRes = 9
};

// Applies a finished grid-wide sum to the solve state (cg.py:90-110).
__device__ __forceinline__ void dot_finalize(CgState* st, int kind, double s, double s2 = 0.0) {
    if (!isfinite(s)) st->nonfinite = 1;
    switch (kind) {
        case kDotUpdate: {  // ||r||^2 after the update, and the test of cg.py:99
            const double norm_r = sqrt(s);
            st->rr = s;
            st->iters += 1;
            if (norm_r < st->atol || norm_r < st->rtol * st->norm_b) st->done = 1;
            break;
        }
        case kDotReorth:  // Polak-Ribiere: s = <r, z>, s2 = <r - r_old, z>
            st->beta = s2 / st->rz;
            st->rz = s;
            break;
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
        case kDotNormB:
            st->norm_b = sqrt(s);
            st->rr = s;
            if (!(s > 0.0)) st->done = 1;
            break;
        case kDotNormBOnly:
            st->norm_b = sqrt(s);
            break;
        case kDotInitRes: {  // cg.py:73-76
            const double norm_r = sqrt(s);
            st->rr = st->rr0 = s;
            if (norm_r < st->atol || norm_r < st->rtol * st->norm_b) st->done = 1;
            break;
        }
        default:
            st->aux = s;
    }
}

// Tile shapes. S = 0 (large grids): 64 x 8 points, two per thread, which
// keeps the halo overhead low. S = 1 (coarse grids, latency-bound): one point
// per thread, so a thread's coefficient loads (25) stay within the loads a
// wave can keep in flight, and a small grid spreads over more workgroups.
#ifndef HM_RESTRICT_CY
#define HM_RESTRICT_CY 4  // coarse rows of a large grid's restriction tile
#endif
template <int S>
struct TileDims { static constexpr int X = 64, Y = 8, CX = 32, CY = HM_RESTRICT_CY; };
template <>
struct TileDims<1> { static constexpr int X = 32, Y = 8, CX = 16, CY = 3; };

// Plain output tiles of OX x OY points starting at (bx*OX, by*OY).
#define HM_TILE_GEOM(S_)                                                        \
    static constexpr int OX = TileDims<S_>::X, OY = TileDims<S_>::Y;            \
    __device__ __forceinline__ int ox0(int bx) const { return bx * OX; }        \
    __device__ __forceinline__ int oy0(int by) const { return by * OY; }

// y = A x, optionally with <x, A x> (the CG SpMV, cg.py:85-90).
template <typename T, int S = 0>
struct OutApply {
    HM_TILE_GEOM(S)
    static constexpr int kProf = 0;
    T* y;
    DotSink dot;
    T* xout = nullptr;  // optional copy of the input at the tile's points
    OutApply<T, 1> small() const { return {y, dot, xout}; }
    static constexpr int kPre = 0;
    __device__ __forceinline__ T preload(int, long long) const { return T(0); }
    __device__ __forceinline__ void point(long long g, int, int, T Ax, T xi, double& acc,
                                          const T*) const {
        y[g] = Ax;
        if (xout) xout[g] = xi;
        acc += (double)xi * (double)Ax;
    }
    __device__ __forceinline__ void outside(int, int) const {}
};

// r = b - A x.
template <typename T, int S = 0>
struct OutResidual {
    HM_TILE_GEOM(S)
    static constexpr int kProf = 3;
    const T* b;
    T* r;
    DotSink dot;
    OutResidual<T, 1> small() const { return {b, r, dot}; }
    static constexpr int kPre = 1;
    __device__ __forceinline__ T preload(int, long long g) const { return b[g]; }
    __device__ __forceinline__ void point(long long g, int, int, T Ax, T, double&,
                                          const T* pv) const {
        r[g] = pv[0] - Ax;
    }
    __device__ __forceinline__ void outside(int, int) const {}
};

// z = x + omega (b - A x) / D (vcycle.py:42), optionally with <b, z>.
template <typename T, int S = 0>
struct OutSmooth {
    HM_TILE_GEOM(S)
    static constexpr int kProf = 2;
    const T* b;
    const T* iD;  // 1 / D
    T omega;
    T* z;
    DotSink dot;
    OutSmooth<T, 1> small() const { return {b, iD, omega, z, dot}; }
    static constexpr int kPre = 2;
    __device__ __forceinline__ T preload(int j, long long g) const { return j == 0 ? b[g] : iD[g]; }
    __device__ __forceinline__ void point(long long g, int, int, T Ax, T xi, double& acc,
                                          const T* pv) const {
        const T bg = pv[0];
        const T zg = vc_smooth(xi, bg, Ax, pv[1], omega);
        z[g] = zg;
        acc += (double)bg * (double)zg;
    }
    __device__ __forceinline__ void outside(int, int) const {}
};

// Residual b - A x on the fine points of a coarse tile, restricted with the
// full-weighting stencil to bc (vcycle.py:77-80).
template <typename T, int S = 0>
struct OutRestrict {
    static constexpr int kProf = 1;
    static constexpr int CX = TileDims<S>::CX, CY = TileDims<S>::CY;
    static constexpr int OX = 2 * CX + 1, OY = 2 * CY + 1;
    const T* b;
    T* bc;
    int w2, h2;
    T* sres;  // set by the engine: LDS [OY][OX]
    DotSink dot;
    OutRestrict<T, 1> small() const { return {b, bc, w2, h2, sres, dot}; }
    __device__ __forceinline__ int ox0(int bx) const { return 2 * CX * bx - 1; }
    __device__ __forceinline__ int oy0(int by) const { return 2 * CY * by - 1; }
    static constexpr int kPre = 1;
    __device__ __forceinline__ T preload(int, long long g) const { return b[g]; }
    __device__ __forceinline__ void point(long long, int lx, int ly, T Ax, T, double&,
                                          const T* pv) const {
        sres[ly * OX + lx] = pv[0] - Ax;
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
            bc[(long long)Y * w2 + X] = vc_restrict9(r0, r1, r2);
        }
    }
};

template <class Out>
struct IsRestrict { static constexpr bool value = false; };
template <typename T, int S>
struct IsRestrict<OutRestrict<T, S>> { static constexpr bool value = true; };

// Large-grid restriction tiles hand out their (2CX+1) x (2CY+1) residual
// points in an order aligned with the grid's 64-point rows: first the 2CX x
// 2CY block the tile owns (local (1, 1) on, two rows per wavefront pair),
// then the row above it, then the column left of both. A row-major walk of
// the region starts one point before a 64-point boundary, and each
// wavefront's coefficient loads then span three cache lines instead of two.
template <class Out>
struct AlignedRestrict { static constexpr bool value = false; };
template <typename T>
struct AlignedRestrict<OutRestrict<T, 0>> { static constexpr bool value = true; };

// Local coordinates (in the output region) of the tile's t-th point.
template <class Out>
__device__ __forceinline__ void out_local(int t, int& lx, int& ly) {
    if constexpr (AlignedRestrict<Out>::value) {
        constexpr int BX = 2 * Out::CX, BY = 2 * Out::CY;
        static_assert(Out::OX == BX + 1 && Out::OY == BY + 1, "restriction region");
        if (t < BX * BY) {
            const int q = t / BX;
            ly = 1 + q;
            lx = 1 + (t - q * BX);
        } else if (t < BX * BY + BX) {
            ly = 0;
            lx = 1 + (t - BX * BY);
        } else {
            ly = t - BX * BY - BX;
            lx = 0;
        }
    } else {
        ly = t / Out::OX;
        lx = t - ly * Out::OX;
    }
}

// -------------------------------------------------------------------- kernel

// Workgroups are dispatched round-robin over the XCDs (workgroup b runs on
// XCD b % kXcd), and each XCD has its own L2. Tile t (row-major over the
// grid's tiles) is given to the workgroup that makes XCD j process the j-th
// contiguous band of tiles in raster order, so the halo rows and columns a
// tile shares with its neighbours are read while they are in the same L2.
constexpr unsigned kXcd = 8;

struct TileMap {
    unsigned tiles_x, ntiles;
    __device__ __forceinline__ void tile(unsigned b, int& bx, int& by) const {
        const unsigned q = ntiles / kXcd, r = ntiles % kXcd;
        const unsigned j = b % kXcd, k = b / kXcd;
        const unsigned t = j * q + (j < r ? j : r) + k;
        by = (int)(t / tiles_x);
        bx = (int)(t - (unsigned)by * tiles_x);
    }
};

// XCD-banded tile of workgroup b in a launch of ntiles tiles (see above).
__device__ __forceinline__ unsigned xcd_tile(unsigned b, unsigned ntiles) {
    const unsigned q = ntiles / kXcd, r = ntiles % kXcd;
    const unsigned j = b % kXcd, k = b / kXcd;
    return j * q + (j < r ? j : r) + k;
}

// One workgroup's tile (bx, by) of an (h, w) grid: stage the input tile and
// its 2-point halo in LDS, apply the operator at the tile's points and hand
// each result to the output; a reduction writes its partial at `slot`.
// `done` (the solve's freeze flag) is tested after the loads are issued, so
// its latency overlaps theirs instead of adding to it.
template <typename T, class Op, class In, class Out>
__device__ __forceinline__ void tile_body(const Op& op, const In& in, const Out& out, int h,
                                          int w, bool done, int bx, int by, unsigned slot) {
    constexpr int OX = Out::OX, OY = Out::OY;
    constexpr int RX = OX + 4, RY = OY + 4;
    constexpr int NO = (OX * OY + kBlock - 1) / kBlock;
    constexpr int NP = Out::kPre > 0 ? Out::kPre : 1;
    __shared__ T sx[RY][RX];
    __shared__ typename Op::template Shared<OX, OY> sh;
    __shared__ T sres[IsRestrict<Out>::value ? OX * OY : 1];
    const int ox0 = out.ox0(bx), oy0 = out.oy0(by);
    // Operators that prefetch per output point walk the region row-major.
    static_assert(!AlignedRestrict<Out>::value || !Op::kPointPrefetch,
                  "aligned restriction tiles need operators without per-point prefetch");
    typename Op::template Prefetch<OX, OY> pf;
    op.template prefetch<OX, OY>(pf, ox0, oy0);
    // Per-point output operands (b, D), loaded before the barrier.
    T pre[NO][NP];
    if constexpr (Out::kPre > 0) {
#pragma unroll
        for (int k = 0; k < NO; ++k) {
            const int t = threadIdx.x + k * kBlock;
            int lx, ly;
            out_local<Out>(t, lx, ly);
            const int gx = ox0 + lx, gy = oy0 + ly;
            const bool in = t < OX * OY && gx >= 0 && gx < w && gy >= 0 && gy < h;
            const long long g = in ? (long long)gy * w + gx : 0;
#pragma unroll
            for (int j = 0; j < Out::kPre; ++j) pre[k][j] = in ? out.preload(j, g) : T(0);
        }
    }

    __shared__ typename InShared<In, RX, RY>::type sin;
    static_assert(!InStaged<In>::value || (!IsRestrict<Out>::value && OX % 2 == 0 && OY % 2 == 0),
                  "staged inputs assume plain tiles starting at even coordinates");
    if constexpr (InStaged<In>::value) {
        in.template stage<RX, RY>(sin, ox0, oy0);
        __syncthreads();
    }
    constexpr int NR = (RX * RY + kBlock - 1) / kBlock;
#pragma unroll
    for (int k = 0; k < NR; ++k) {
        const int t = threadIdx.x + k * kBlock;
        if (t < RX * RY) {
            const int ly = t / RX, lx = t - ly * RX;
            const int gx = ox0 - 2 + lx, gy = oy0 - 2 + ly;
            const bool inb = gx >= 0 && gx < w && gy >= 0 && gy < h;
            const long long g = inb ? (long long)gy * w + gx : 0;
            T xv = T(0);
            if constexpr (InStaged<In>::value) {
                if (inb) xv = in.template load_staged<RX, RY>(sin, ox0, oy0, gx, gy, g);
            } else {
                if (inb) xv = in.load(gx, gy, g);
            }
            sx[ly][lx] = xv;
            op.template load_point<OX, OY>(sh, ly, lx, inb, g, xv);
        }
    }
    if (done) return;  // uniform across the workgroup
    __syncthreads();
    op.template prepare<OX, OY>(sh, pf);
    __syncthreads();

    Out o = out;
    if constexpr (IsRestrict<Out>::value) o.sres = sres;
    double acc = 0.0;
#pragma unroll
    for (int k = 0; k < NO; ++k) {
        const int t = threadIdx.x + k * kBlock;
        if (t >= OX * OY) continue;
        int lx, ly;
        out_local<Out>(t, lx, ly);
        const int gx = ox0 + lx, gy = oy0 + ly;
        if (gx < 0 || gx >= w || gy < 0 || gy >= h) {
            o.outside(lx, ly);
            continue;
        }
        const long long g = (long long)gy * w + gx;
        const T xi = sx[ly + 2][lx + 2];
        const T Ax = op.template apply<OX, OY>(sh, pf, sx, k, lx, ly, xi, g);
        o.point(g, lx, ly, Ax, xi, acc, pre[k]);
    }
    if constexpr (IsRestrict<Out>::value) {
        __syncthreads();
        o.restrict_tile(bx, by);
    }
    if (o.dot.st != nullptr) {
        double v[1] = {acc};
        block_partials_at<kBlock, 1>(v, o.dot.red, slot);
    }
}

// kLevel only names the grid class of a launch in traces (0 the finest grid,
// 1 the first Galerkin grid, 2 the coarser ones); the code is the same.
template <typename T, class Op, class In, class Out, int kLevel = 0>
__global__ __launch_bounds__(kBlock) void tile_kernel(Op op, In in, Out out, int h, int w,
                                                      const CgState* guard, TileMap map) {
    const bool done = solve_done(guard);
    int bx, by;
    map.tile(blockIdx.x, bx, by);
    tile_body<T, Op, In, Out>(op, in, out, h, w, done, bx, by, blockIdx.x);
}

// Stage two of every reduction: one 1024-thread workgroup sums nblocks
// partials per value in a fixed order and applies the result to the solve
// state. All loads of a thread are issued before the first add.
constexpr int kFinBlock = 1024;

// Fixed-order sums of cnt[k] partials at src[k] for k < N, valid in thread 0.
template <int N>
__device__ __forceinline__ void finalize_sums(const double* const (&src)[N],
                                              const unsigned (&cnt)[N], double (&acc)[N]) {
    constexpr int U = 4;
    unsigned most = 0;
#pragma unroll
    for (int k = 0; k < N; ++k) most = cnt[k] > most ? cnt[k] : most;
    double part[N][U];
#pragma unroll
    for (int k = 0; k < N; ++k)
#pragma unroll
        for (int u = 0; u < U; ++u) part[k][u] = 0.0;
    for (unsigned i0 = 0; i0 < most; i0 += kFinBlock * U) {
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const unsigned i = i0 + u * kFinBlock + threadIdx.x;
#pragma unroll
            for (int k = 0; k < N; ++k)
                if (i < cnt[k]) part[k][u] += src[k][i];
        }
    }
#pragma unroll
    for (int k = 0; k < N; ++k) {
        double t = 0.0;
#pragma unroll
        for (int u = 0; u < U; ++u) t += part[k][u];
        acc[k] = t;
    }
    block_sum<kFinBlock, N>(acc);
}

// Applies finished sums to the solve state (thread 0 only). kDotUpdateRz: the
// residual norm of the CG update (acc[0]) and <r, z> of the preconditioner
// that follows it (acc[N-1]): the convergence test first, then beta only for
// an unconverged solve.
template <int N>
__device__ __forceinline__ void finalize_apply(CgState* st, int kind, const double (&acc)[N]) {
    if (kind == kDotUpdateRz) {
        dot_finalize(st, kDotUpdate, acc[0]);
        if (!st->done) dot_finalize(st, kDotRzUpdate, acc[N - 1]);
    } else {
        dot_finalize(st, kind, acc[0], N > 1 ? acc[N - 1] : 0.0);
    }
}

template <int N>
__global__ __launch_bounds__(kFinBlock) void k_finalize(Reducer red, unsigned nblocks,
                                                        CgState* st, int kind,
                                                        const CgState* guard) {
    const bool done = solve_done(guard);
    const double* src[N];
    unsigned cnt[N];
#pragma unroll
    for (int k = 0; k < N; ++k) {
        src[k] = red.partials + (size_t)k * red.max_blocks;
        cnt[k] = nblocks;
    }
    if (kind == kDotUpdateRz) {  // slot 2: the update's partials; slot 0: <r, z>
        src[0] = red.partials + 2 * (size_t)red.max_blocks;
        cnt[0] = red.upd_blocks;
        if (N > 1) {
            src[N - 1] = red.partials;
            cnt[N - 1] = nblocks;
        }
    }
    if (done) return;  // uniform across the workgroup
    double acc[N];
    finalize_sums<N>(src, cnt, acc);
    if (threadIdx.x == 0) finalize_apply<N>(st, kind, acc);
}

template <class Out>
inline TileMap tile_map(const Out&, int h, int w, int h2, int w2) {
    if constexpr (IsRestrict<Out>::value) {
        const unsigned tx = div_up(w2, Out::CX), ty = div_up(h2, Out::CY);
        return TileMap{tx, tx * ty};
    } else {
        const unsigned tx = div_up(w, Out::OX), ty = div_up(h, Out::OY);
        return TileMap{tx, tx * ty};
    }
}

}  // namespace hm
