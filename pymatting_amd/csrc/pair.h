// Two consecutive coarse levels of a V-cycle in one launch.
//
// On the coarse grids of the hierarchy (below ~2^18 points) a level costs
// what its launch costs -- a kernel boundary and one trip to memory -- not
// its bytes. The two kernels here cover two levels each, with one pre- and
// one post-smoothing sweep (vcycle.py:46-100 with num_pre_iter =
// num_post_iter = 1):
//  - k_vc_pair_down, the descent: on level a, the first damped-Jacobi
//    iterate z = omega b / D, the residual b - A z and its full-weighting
//    restriction b_{a+1}; the same on level a+1, giving b_{a+2};
//  - k_vc_pair_up, the ascent: on level a+1, the first iterate plus the
//    prolongated correction of level a+2 and one damped-Jacobi sweep, giving
//    z_{a+1}; the same on level a with z_{a+1} as the correction, giving z_a.
// A workgroup owns a tile of the level it writes and recomputes the level in
// between (and the input level's halo) on the region its stencils, the
// restriction and the prolongation reach, from the same inputs and with the
// same arithmetic as the one-level tile kernels (the vc_* helpers and the
// 25-tap fused multiply-add chain in slot order of engine.h). A pair
// therefore gives bit for bit what two one-level launches give. Points
// outside a grid hold zeros, as in the one-level tiles.
#pragma once

#include <vector>

#include "vcycle.h"

namespace hm {

constexpr int cdiv(int a, int b) { return (a + b - 1) / b; }

template <typename T>
struct PairLevel {
    int h, w;
    const T* coef;  // [13][n] symmetric storage
    const T* iD;    // 1 / diagonal
    T* b;           // right-hand side (written by the descent for level a+1)
};

// Regions of the descent for an output tile of CX x CY points of level a+2
// at (X0, Y0), each in its own level's coordinates:
//   R1 = [2 X0 - 1, 2 X0 + 2 CX - 1]: level a+1 residuals under the tile's
//        restriction stencils;
//   R2 = R1 grown by 2 (the stencil reach): level a+1 right-hand side and
//        first iterate;
//   R3 = [2 (2 X0 - 3) - 1, 2 (2 X0 + 2 CX + 1) + 1]: level a residuals under
//        the restriction stencils of R2;
//   R4 = R3 grown by 2: level a first iterate.
template <int CX, int CY>
struct PairDownDims {
    static constexpr int R1X = 2 * CX + 1, R1Y = 2 * CY + 1;
    static constexpr int R2X = R1X + 4, R2Y = R1Y + 4;
    static constexpr int R3X = 2 * R2X + 1, R3Y = 2 * R2Y + 1;
    static constexpr int R4X = R3X + 4, R4Y = R3Y + 4;
    static constexpr int N1 = R1X * R1Y, N2 = R2X * R2Y, N3 = R3X * R3Y, N4 = R4X * R4Y;
};

// NPF: rounds of level-a residual points whose 25 coefficients are loaded
// into registers before the first wait (the rest are loaded at use).
template <typename T, int CX, int CY, int NPF, int BS>
__global__ __launch_bounds__(BS) void k_vc_pair_down(PairLevel<T> A, PairLevel<T> B,
                                                             T* __restrict__ b2, int h2, int w2,
                                                             T omega, const CgState* guard,
                                                             TileMap map) {
    using D = PairDownDims<CX, CY>;
    constexpr int K3 = cdiv(D::N3, BS), K4 = cdiv(D::N4, BS);
    constexpr int K2 = cdiv(D::N2, BS), KC = cdiv(kSymSlots * D::N2, BS);
    static_assert(NPF >= 0 && NPF <= K3, "prefetch rounds");
    __shared__ T s_z4[D::N4];               // level a first iterate on R4
    __shared__ T s_r3[D::N3];               // level a residuals on R3
    __shared__ T s_c2[kSymSlots * D::N2];   // level a+1 stencil on R2
    __shared__ T s_b2[D::N2];               // level a+1 right-hand side on R2
    __shared__ T s_z2[D::N2];               // level a+1 first iterate on R2
    __shared__ T s_r1[D::N1];               // level a+1 residuals on R1
    const bool done = solve_done(guard);
    int bx, by;
    map.tile(blockIdx.x, bx, by);
    const int X0 = CX * bx, Y0 = CY * by;
    const int x3 = 4 * X0 - 7, y3 = 4 * Y0 - 7;  // R3 origin (level a)
    const int x2 = 2 * X0 - 3, y2 = 2 * Y0 - 3;  // R2 origin (level a+1)
    const int na = A.h * A.w, nb = B.h * B.w;
    const int tid = threadIdx.x;

    // Every global load of the tile, issued before the first wait.
    T c3[NPF > 0 ? NPF : 1][25];
    T b3[K3];
#pragma unroll
    for (int k = 0; k < K3; ++k) {
        const int t = tid + k * BS;
        const int ly = t / D::R3X, lx = t - ly * D::R3X;
        const int gx = x3 + lx, gy = y3 + ly;
        const bool in = t < D::N3 && gx >= 0 && gx < A.w && gy >= 0 && gy < A.h;
        b3[k] = in ? A.b[gy * A.w + gx] : T(0);
    }
#pragma unroll
    for (int k = 0; k < NPF; ++k) {
        const int t = tid + k * BS;
        const int ly = t / D::R3X, lx = t - ly * D::R3X;
        const int gx = x3 + lx, gy = y3 + ly;
        const bool in = t < D::N3 && gx >= 0 && gx < A.w && gy >= 0 && gy < A.h;
        const int g = in ? gy * A.w + gx : 0;
#pragma unroll
        for (int s = 0; s < 25; ++s) c3[k][s] = in ? sym_coef32(A.coef, na, A.w, g, s) : T(0);
    }
    T z4[K4];
#pragma unroll
    for (int k = 0; k < K4; ++k) {
        const int t = tid + k * BS;
        const int ly = t / D::R4X, lx = t - ly * D::R4X;
        const int gx = x3 - 2 + lx, gy = y3 - 2 + ly;
        const bool in = t < D::N4 && gx >= 0 && gx < A.w && gy >= 0 && gy < A.h;
        const int g = in ? gy * A.w + gx : 0;
        z4[k] = in ? vc_jac0(omega, A.b[g], A.iD[g]) : T(0);
    }
    T c2[KC];
#pragma unroll
    for (int k = 0; k < KC; ++k) {
        const int t = tid + k * BS;
        const int q = t / D::N2, p = t - q * D::N2;
        const int py = p / D::R2X, px = p - py * D::R2X;
        const int gx = x2 + px, gy = y2 + py;
        const bool in = t < kSymSlots * D::N2 && gx >= 0 && gx < B.w && gy >= 0 && gy < B.h;
        c2[k] = in ? B.coef[(size_t)q * nb + gy * B.w + gx] : T(0);
    }
    T id2[K2];
#pragma unroll
    for (int k = 0; k < K2; ++k) {
        const int t = tid + k * BS;
        const int py = t / D::R2X, px = t - py * D::R2X;
        const int gx = x2 + px, gy = y2 + py;
        const bool in = t < D::N2 && gx >= 0 && gx < B.w && gy >= 0 && gy < B.h;
        id2[k] = in ? B.iD[gy * B.w + gx] : T(0);
    }
    if (done) return;  // uniform across the workgroup
#pragma unroll
    for (int k = 0; k < K4; ++k) {
        const int t = tid + k * BS;
        if (t < D::N4) s_z4[t] = z4[k];
    }
#pragma unroll
    for (int k = 0; k < KC; ++k) {
        const int t = tid + k * BS;
        if (t < kSymSlots * D::N2) s_c2[t] = c2[k];
    }
    __syncthreads();

    // Level a residuals b - A z on R3 (zero outside the grid).
#pragma unroll
    for (int k = 0; k < K3; ++k) {
        const int t = tid + k * BS;
        if (t >= D::N3) continue;
        const int ly = t / D::R3X, lx = t - ly * D::R3X;
        const int gx = x3 + lx, gy = y3 + ly;
        T res = T(0);
        if (gx >= 0 && gx < A.w && gy >= 0 && gy < A.h) {
            const auto zs = [&](int s) { return s_z4[(ly + s / 5) * D::R4X + lx + s % 5]; };
            T Az;
            if (k < NPF) {
                Az = vc_stencil<T>([&](int s) { return c3[k < NPF ? k : 0][s]; }, zs);
            } else {
                const int g = gy * A.w + gx;
                Az = vc_stencil<T>([&](int s) { return sym_coef32(A.coef, na, A.w, g, s); }, zs);
            }
            res = b3[k] - Az;
        }
        s_r3[t] = res;
    }
    __syncthreads();

    // Level a+1 right-hand side (restriction) and first iterate on R2; the
    // tile writes the right-hand side of the level a+1 points it owns:
    // [2 X0, 2 X0 + 2 CX) per axis, the last tile of a row or column also the
    // odd point left over at the grid's end.
    const unsigned tiles_x = map.tiles_x, tiles_y = map.ntiles / map.tiles_x;
    const int ox_hi = (unsigned)bx + 1 == tiles_x ? B.w : 2 * X0 + 2 * CX;
    const int oy_hi = (unsigned)by + 1 == tiles_y ? B.h : 2 * Y0 + 2 * CY;
#pragma unroll
    for (int k = 0; k < K2; ++k) {
        const int t = tid + k * BS;
        if (t >= D::N2) continue;
        const int py = t / D::R2X, px = t - py * D::R2X;
        const int gx = x2 + px, gy = y2 + py;
        T bb = T(0), zz = T(0);
        if (gx >= 0 && gx < B.w && gy >= 0 && gy < B.h) {
            const T* r = s_r3 + (2 * py + 1) * D::R3X + 2 * px + 1;
            bb = vc_restrict9(r - D::R3X, r, r + D::R3X);
            zz = vc_jac0(omega, bb, id2[k]);
            if (gx >= 2 * X0 && gx < ox_hi && gy >= 2 * Y0 && gy < oy_hi)
                B.b[gy * B.w + gx] = bb;
        }
        s_b2[t] = bb;
        s_z2[t] = zz;
    }
    __syncthreads();

    // Level a+1 residuals on R1; the stencil's slots s < 12 are those of the
    // neighbour's mirrored slot (symmetric storage), all within R2.
    for (int t = tid; t < D::N1; t += BS) {
        const int ly = t / D::R1X, lx = t - ly * D::R1X;
        const int gx = x2 + 2 + lx, gy = y2 + 2 + ly;
        T res = T(0);
        if (gx >= 0 && gx < B.w && gy >= 0 && gy < B.h) {
            const int p = (ly + 2) * D::R2X + lx + 2;
            const T Az = vc_stencil<T>(
                [&](int s) {
                    return s >= 12 ? s_c2[(s - 12) * D::N2 + p]
                                   : s_c2[(12 - s) * D::N2 + p + (s / 5 - 2) * D::R2X + (s % 5 - 2)];
                },
                [&](int s) { return s_z2[(ly + s / 5) * D::R2X + lx + s % 5]; });
            res = s_b2[p] - Az;
        }
        s_r1[t] = res;
    }
    __syncthreads();

    // Level a+2 right-hand side at the tile's points.
    for (int t = tid; t < CX * CY; t += BS) {
        const int cy = t / CX, cx = t - cy * CX;
        const int X = X0 + cx, Y = Y0 + cy;
        if (X >= w2 || Y >= h2) continue;
        const T* r = s_r1 + (2 * cy + 1) * D::R1X + 2 * cx + 1;
        b2[Y * w2 + X] = vc_restrict9(r - D::R1X, r, r + D::R1X);
    }
}

// The ascent as a chain of NL levels a, a+1, ..., a+NL-1 in one launch: the
// correction e of level a+NL is prolongated and smoothed level by level up to
// level a. A workgroup owns an output tile of OX x OY points of level a at
// (ox0, oy0) (even); in each level's own coordinates it works on
//   T2  = the tile grown by 2 (level a): the iterate before level a's sweep;
//   Q_j = the level a+j points the prolongation reads from the finer level's
//         iterate region (j >= 1): the corrections after level a+j's sweep;
//   Q2_j = Q_j grown by 2: the iterate before level a+j's sweep;
//   E   = the level a+NL corrections the prolongation reads on Q2_{NL-1}.
// Going down the chain each region halves and grows by a few points, so the
// levels below a add only a fraction of level a's work.
constexpr int kChainMax = 5;  // levels per launch, at most

struct ChainSizes {
    int qx[kChainMax], qy[kChainMax], q2x[kChainMax], q2y[kChainMax];
    int ex, ey;
};

// Region sizes (upper bounds over the tile's position; the first coarse
// region's origin is even, so its size is exact).
template <int OX, int OY, int NL>
constexpr ChainSizes chain_sizes() {
    ChainSizes c{};
    int wx = OX + 4, wy = OY + 4;  // iterate region of the finer level
    for (int j = 1; j < NL; ++j) {
        c.qx[j] = j == 1 ? OX / 2 + 3 : (wx + 1) / 2 + 2;
        c.qy[j] = j == 1 ? OY / 2 + 3 : (wy + 1) / 2 + 2;
        c.q2x[j] = c.qx[j] + 4;
        c.q2y[j] = c.qy[j] + 4;
        wx = c.q2x[j];
        wy = c.q2y[j];
    }
    c.ex = (wx + 1) / 2 + 2;
    c.ey = (wy + 1) / 2 + 2;
    return c;
}

// LDS layout, staged part first (in the order the staging loop fills it):
// level a's b and 1/D on T2; per level j >= 1 its 13 stencil planes, b and
// 1/D on Q2_j; the correction on E. Then the computed part: level a's
// iterate on T2, per level j its iterate on Q2_j and correction on Q_j.
struct ChainLds {
    int nt2, ba, ida, c[kChainMax], b[kChainMax], id[kChainMax], e, staged;
    int ya, y[kChainMax], z[kChainMax], total;
};

template <int OX, int OY, int NL>
constexpr ChainLds chain_lds() {
    constexpr ChainSizes S = chain_sizes<OX, OY, NL>();
    ChainLds o{};
    o.nt2 = (OX + 4) * (OY + 4);
    int off = 0;
    o.ba = off;
    off += o.nt2;
    o.ida = off;
    off += o.nt2;
    for (int j = 1; j < NL; ++j) {
        const int n2 = S.q2x[j] * S.q2y[j];
        o.c[j] = off;
        off += kSymSlots * n2;
        o.b[j] = off;
        off += n2;
        o.id[j] = off;
        off += n2;
    }
    o.e = off;
    off += S.ex * S.ey;
    o.staged = off;
    o.ya = off;
    off += o.nt2;
    for (int j = 1; j < NL; ++j) {
        o.y[j] = off;
        off += S.q2x[j] * S.q2y[j];
        o.z[j] = off;
        off += S.qx[j] * S.qy[j];
    }
    o.total = off;
    return o;
}

template <typename T>
struct ChainArgs {
    PairLevel<T> lv[kChainMax];  // lv[j]: level a+j
    const T* e;                  // correction of level a+NL
    int he, we;
    T omega;
    T* z;  // post-smoothed iterate of level a
};

// PFA: the level a stencil of the thread's output point is loaded into
// registers before the first wait (else at use, in the last phase).
template <typename T, int OX, int OY, int NL, bool PFA, int BS>
__global__ __launch_bounds__(BS) void k_vc_up_chain(ChainArgs<T> P, const CgState* guard,
                                                    TileMap map) {
    static_assert(NL >= 2 && NL <= kChainMax, "chain length");
    static_assert(OX % 2 == 0 && OY % 2 == 0, "tiles start at even coordinates");
    static_assert(OX * OY <= BS, "one output point per thread");
    constexpr ChainSizes S = chain_sizes<OX, OY, NL>();
    constexpr ChainLds O = chain_lds<OX, OY, NL>();
    constexpr int TX = OX + 4, TY = OY + 4, NT2 = TX * TY;
    constexpr int KS = cdiv(O.staged, BS);
    __shared__ T lds[O.total];
    const bool done = solve_done(guard);
    int bx, by;
    map.tile(blockIdx.x, bx, by);
    const int ox0 = OX * bx, oy0 = OY * by;
    const int tx0 = ox0 - 2, ty0 = oy0 - 2;  // T2 origin (level a)
    // Region origins of the coarser levels (>> floors).
    int qx0[kChainMax], qy0[kChainMax];
    {
        int px = tx0, py = ty0;
#pragma unroll
        for (int j = 1; j < NL; ++j) {
            qx0[j] = px >> 1;
            qy0[j] = py >> 1;
            px = qx0[j] - 2;
            py = qy0[j] - 2;
        }
        qx0[0] = px >> 1;  // E origin, kept in slot 0
        qy0[0] = py >> 1;
    }
    const int tid = threadIdx.x;
    const PairLevel<T>& A = P.lv[0];
    const int na = A.h * A.w;

    // Every global load of the tile, issued before the first wait.
    const int oly = tid / OX, olx = tid - oly * OX;
    const int ogx = ox0 + olx, ogy = oy0 + oly;
    const bool own = tid < OX * OY && ogx < A.w && ogy < A.h;
    const int og = own ? ogy * A.w + ogx : 0;
    T ca[PFA ? 25 : 1];
    if constexpr (PFA) {
#pragma unroll
        for (int s = 0; s < 25; ++s) ca[s] = own ? sym_coef32(A.coef, na, A.w, og, s) : T(0);
    }
    T sv[KS];
#pragma unroll
    for (int k = 0; k < KS; ++k) {
        int t = tid + k * BS;
        T v = T(0);
        if (t < 2 * NT2) {  // level a: b, then 1/D, on T2
            const int q = t >= NT2, p = t - q * NT2;
            const int ly = p / TX, lx = p - ly * TX;
            const int gx = tx0 + lx, gy = ty0 + ly;
            if (gx >= 0 && gx < A.w && gy >= 0 && gy < A.h) v = (q ? A.iD : A.b)[gy * A.w + gx];
        } else {
            t -= 2 * NT2;
            bool found = false;
#pragma unroll
            for (int j = 1; j < NL; ++j) {
                const int n2 = S.q2x[j] * S.q2y[j];
                if (!found && t < (kSymSlots + 2) * n2) {
                    found = true;
                    const int q = t / n2, p = t - q * n2;  // plane 0..12, 13: b, 14: 1/D
                    const int ly = p / S.q2x[j], lx = p - ly * S.q2x[j];
                    const int gx = qx0[j] - 2 + lx, gy = qy0[j] - 2 + ly;
                    const PairLevel<T>& L = P.lv[j];
                    if (gx >= 0 && gx < L.w && gy >= 0 && gy < L.h) {
                        const int g = gy * L.w + gx;
                        v = q < kSymSlots ? L.coef[(size_t)q * L.h * L.w + g]
                                          : (q == kSymSlots ? L.b[g] : L.iD[g]);
                    }
                } else if (!found) {
                    t -= (kSymSlots + 2) * n2;
                }
            }
            if (!found && t < S.ex * S.ey) {  // the correction on E
                const int ly = t / S.ex, lx = t - ly * S.ex;
                const int gx = qx0[0] + lx, gy = qy0[0] + ly;
                if (gx >= 0 && gx < P.we && gy >= 0 && gy < P.he) v = P.e[gy * P.we + gx];
            }
        }
        sv[k] = v;
    }
    if (done) return;  // uniform across the workgroup
#pragma unroll
    for (int k = 0; k < KS; ++k) {
        const int t = tid + k * BS;
        if (t < O.staged) lds[t] = sv[k];
    }
    __syncthreads();

    // Coarsest to finest: level a+j's iterate omega b / D + P (correction of
    // level a+j+1) on Q2_j, then its sweep on Q_j.
#pragma unroll
    for (int j = NL - 1; j >= 1; --j) {
        const PairLevel<T>& L = P.lv[j];
        const int n2 = S.q2x[j] * S.q2y[j];
        const T* cz = j == NL - 1 ? lds + O.e : lds + O.z[j + 1 < NL ? j + 1 : j];
        const int czw = j == NL - 1 ? S.ex : S.qx[j + 1 < NL ? j + 1 : j];
        const int czx = j == NL - 1 ? qx0[0] : qx0[j + 1 < NL ? j + 1 : j];
        const int czy = j == NL - 1 ? qy0[0] : qy0[j + 1 < NL ? j + 1 : j];
        const int x2 = qx0[j] - 2, y2 = qy0[j] - 2;
        T* yj = lds + O.y[j];
        const T* bj = lds + O.b[j];
        const T* idj = lds + O.id[j];
        for (int t = tid; t < n2; t += BS) {
            const int ly = t / S.q2x[j], lx = t - ly * S.q2x[j];
            const int gx = x2 + lx, gy = y2 + ly;
            T y = T(0);
            if (gx >= 0 && gx < L.w && gy >= 0 && gy < L.h) {
                const bool ox = gx & 1, oy = gy & 1;
                const T* e = cz + ((gy >> 1) - czy) * czw + (gx >> 1) - czx;
                y = vc_prolong(vc_jac0(P.omega, bj[t], idj[t]), ox, oy, e[0], ox ? e[1] : T(0),
                               oy ? e[czw] : T(0), (ox && oy) ? e[czw + 1] : T(0));
            }
            yj[t] = y;
        }
        __syncthreads();
        const T* cj = lds + O.c[j];
        T* zj = lds + O.z[j];
        const int qw = S.qx[j], q2w = S.q2x[j];
        for (int t = tid; t < S.qx[j] * S.qy[j]; t += BS) {
            const int qy = t / qw, qx = t - qy * qw;
            const int gx = qx0[j] + qx, gy = qy0[j] + qy;
            T zz = T(0);
            if (gx >= 0 && gx < L.w && gy >= 0 && gy < L.h) {
                const int p = (qy + 2) * q2w + qx + 2;
                const T Ay = vc_stencil<T>(
                    [&](int s) {
                        return s >= 12 ? cj[(s - 12) * n2 + p]
                                       : cj[(12 - s) * n2 + p + (s / 5 - 2) * q2w + (s % 5 - 2)];
                    },
                    [&](int s) { return yj[(qy + s / 5) * q2w + qx + s % 5]; });
                zz = vc_smooth(yj[p], bj[p], Ay, idj[p], P.omega);
            }
            zj[t] = zz;
        }
        __syncthreads();
    }

    // Level a: iterate on T2 from level a+1's correction, then the sweep at
    // the tile's points.
    {
        T* ya = lds + O.ya;
        const T* z1 = lds + O.z[1];
        const int qw = S.qx[1];
        for (int t = tid; t < NT2; t += BS) {
            const int ly = t / TX, lx = t - ly * TX;
            const int gx = tx0 + lx, gy = ty0 + ly;
            T y = T(0);
            if (gx >= 0 && gx < A.w && gy >= 0 && gy < A.h) {
                const bool ox = gx & 1, oy = gy & 1;
                const T* e = z1 + ((gy >> 1) - qy0[1]) * qw + (gx >> 1) - qx0[1];
                y = vc_prolong(vc_jac0(P.omega, lds[O.ba + t], lds[O.ida + t]), ox, oy, e[0],
                               ox ? e[1] : T(0), oy ? e[qw] : T(0),
                               (ox && oy) ? e[qw + 1] : T(0));
            }
            ya[t] = y;
        }
        __syncthreads();
        if (own) {
            const int p = (oly + 2) * TX + olx + 2;
            const auto ys = [&](int s) { return ya[(oly + s / 5) * TX + olx + s % 5]; };
            T Ay;
            if constexpr (PFA)
                Ay = vc_stencil<T>([&](int s) { return ca[s]; }, ys);
            else
                Ay = vc_stencil<T>([&](int s) { return sym_coef32(A.coef, na, A.w, og, s); }, ys);
            P.z[og] = vc_smooth(ya[p], lds[O.ba + p], Ay, lds[O.ida + p], P.omega);
        }
    }
}

// Tile shapes: the descent pair writes 8 x 8 points of level a+2 (a 43 x 43
// region of level a residuals per workgroup); the ascent chain writes
// 32 x 16 points of level a (one per thread of a 1024-thread workgroup, half
// of the workgroup idle in the last phase only).
#ifndef HM_PAIR_DCX
#define HM_PAIR_DCX 8
#endif
#ifndef HM_PAIR_DCY
#define HM_PAIR_DCY 8
#endif
#ifndef HM_PAIR_DBS
#define HM_PAIR_DBS 1024
#endif
#ifndef HM_CHAIN_OX
#define HM_CHAIN_OX 32
#endif
#ifndef HM_CHAIN_OY
#define HM_CHAIN_OY 16
#endif
#ifndef HM_CHAIN_BS
#define HM_CHAIN_BS 1024
#endif
constexpr int kPairDownCX = HM_PAIR_DCX, kPairDownCY = HM_PAIR_DCY, kPairDownBlock = HM_PAIR_DBS;
constexpr int kChainOX = HM_CHAIN_OX, kChainOY = HM_CHAIN_OY, kChainBlock = HM_CHAIN_BS;

// Which coarse levels run together: levels from HM_CHAIN_FROM on with fewer
// than 2^HM_CHAIN_MAX_LOG2 points, before the tail (or the coarsest grid), of
// a cycle with one pre- and one post-smoothing sweep; up to HM_CHAIN_LEN
// levels per ascent launch. Measured at 1920x1080 (levels 480x270 .. 60x33):
// one 4-level ascent launch beats two 2-level ones and four one-level ones;
// taking level 1 (960x540) into a chain, or running descent pairs, is slower
// (its halos repeat too much of the finer level's work).
#ifndef HM_CHAIN_FROM
#define HM_CHAIN_FROM 2
#endif
#ifndef HM_CHAIN_MAX_LOG2
#define HM_CHAIN_MAX_LOG2 18
#endif
#ifndef HM_CHAIN_LEN
#define HM_CHAIN_LEN 4
#endif
static_assert(HM_CHAIN_LEN >= 2 && HM_CHAIN_LEN <= kChainMax, "chain length");

// Launch structure bits of hm_vc::pairs (kVcTailMatrix: vcycle.h).
constexpr int kVcChainUp = 1;    // ascent chains
constexpr int kVcPairDown = 2;   // descent pairs

struct VcPlan {
    std::vector<char> down_pair;  // [a]: levels a, a+1 of the descent in one launch
    std::vector<int> up_start;    // [l]: l is the coarsest level of an ascent chain
                                  // starting at up_start[l] (else -1)
};

// desc_end: the descent restricts grids [0, desc_end) (the tail, or the grid
// whose right-hand side the tail matrix reads).
inline VcPlan vc_plan(const hm_vc* vc, size_t tail, size_t desc_end) {
    const size_t L = vc->levels.size();
    VcPlan p;
    p.down_pair.assign(L, 0);
    p.up_start.assign(L, -1);
    if (vc->pairs == 0 || vc->pre != 1 || vc->post != 1) return p;
    size_t l0 = HM_CHAIN_FROM < 1 ? 1 : HM_CHAIN_FROM;
    while (l0 < L && vc->levels[l0].n >= (1LL << HM_CHAIN_MAX_LOG2)) ++l0;
    if (vc->pairs & kVcPairDown)
        for (size_t l = l0; l + 1 < desc_end && l + 2 < L; l += 2) p.down_pair[l] = 1;
    if (vc->pairs & kVcChainUp) {
        for (size_t l = l0; l + 1 < tail && l + 1 < L;) {
            const size_t len = std::min<size_t>(HM_CHAIN_LEN, tail - l);
            if (len < 2) break;
            p.up_start[l + len - 1] = (int)l;
            l += len;
        }
    }
    return p;
}

template <typename T>
PairLevel<T> pair_level(const VcLevel& V) {
    return PairLevel<T>{V.h, V.w, pick<T>(V.A->c64, V.A->c32), pick<T>(V.A->idiag64, V.A->idiag32),
                        static_cast<T*>(V.b)};
}

// Descent of levels a and a+1: b_a -> b_{a+1} (stored), b_{a+2}.
template <typename T>
int launch_pair_down(const hm_vc* vc, size_t a, const CgState* guard, hipStream_t s) {
    const VcLevel& La = vc->levels[a];
    const VcLevel& Lb = vc->levels[a + 1];
    const VcLevel& Lc = vc->levels[a + 2];
    constexpr int CX = kPairDownCX, CY = kPairDownCY;
    const unsigned tx = div_up(Lc.w, CX), ty = div_up(Lc.h, CY);
    const TileMap map{tx, tx * ty};
    ProfScope prof(kProfPairDown, La.n, sizeof(T) == 4, s);
    constexpr int BS = kPairDownBlock;
    constexpr int NPF = sizeof(T) == 4 ? cdiv(PairDownDims<CX, CY>::N3, BS) : 1;
    k_vc_pair_down<T, CX, CY, NPF, BS><<<map.ntiles, BS, 0, s>>>(
        pair_level<T>(La), pair_level<T>(Lb), static_cast<T*>(Lc.b), Lc.h, Lc.w, (T)vc->omega,
        guard, map);
    HM_LAUNCH_CHECK();
    return HM_OK;
}

template <typename T, int NL>
void launch_chain_nl(const ChainArgs<T>& P, const CgState* guard, const TileMap& map,
                     hipStream_t s) {
    // float64 coefficients of the output point would exceed the registers a
    // 1024-thread workgroup has per thread; they are loaded at use.
    k_vc_up_chain<T, kChainOX, kChainOY, NL, sizeof(T) == 4, kChainBlock>
        <<<map.ntiles, kChainBlock, 0, s>>>(P, guard, map);
}

// Ascent of levels a+len-1, ..., a: correction e of level a+len -> z_a.
template <typename T>
int launch_up_chain(const hm_vc* vc, size_t a, size_t len, const T* e, T* z,
                    const CgState* guard, hipStream_t s) {
    if (len < 2 || len > (size_t)kChainMax || a + len >= vc->levels.size()) {
        set_error("internal: V-cycle chain of %zu levels at level %zu", len, a);
        return HM_ERR_ARG;
    }
    ChainArgs<T> P{};
    for (size_t j = 0; j < len; ++j) P.lv[j] = pair_level<T>(vc->levels[a + j]);
    const VcLevel& Le = vc->levels[a + len];
    P.e = e;
    P.he = Le.h;
    P.we = Le.w;
    P.omega = (T)vc->omega;
    P.z = z;
    const VcLevel& La = vc->levels[a];
    const unsigned tx = div_up(La.w, kChainOX), ty = div_up(La.h, kChainOY);
    const TileMap map{tx, tx * ty};
    ProfScope prof(kProfPairUp, La.n, sizeof(T) == 4, s);
    switch (len) {
        case 2: launch_chain_nl<T, 2>(P, guard, map, s); break;
        case 3: launch_chain_nl<T, 3>(P, guard, map, s); break;
        case 4: launch_chain_nl<T, 4>(P, guard, map, s); break;
        default: launch_chain_nl<T, 5>(P, guard, map, s); break;
    }
    HM_LAUNCH_CHECK();
    return HM_OK;
}

}  // namespace hm
