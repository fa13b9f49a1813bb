// Per-edgel reverse pass of the TOED outputs with the edgel set frozen, shared
// by the index-driven backward (toed_backward.hip) and the tile-driven
// backward of a batch plan (toed_detect.hip).  FP64 throughout; one G-lane
// group per edgel (lanes split the window rows, totals by DPP group sums).
//
// Differentiated forward arithmetic: the nine responses at a pixel
// (cpu_toed.cpp:194-221), octant choice and neighbour interpolation
// (:415-462), parabola and subpixel position (:465-489), strength (:481-485),
// third-order orientation (:227-232).
#pragma once

#include <hip/hip_runtime.h>

#include "toed_dpp.hpp"

namespace toedadj {

constexpr int G = 8;    // default lanes per edgel (window rows split over the group)

// response k (fx, fy, fxx, fxy, fyy, fxxy, fxyy, fxxx, fyyy) -> x (column) and
// y (row) filter family (0 G, 1 G', 2 G'', 3 G''')
__host__ __device__ constexpr int XF(int k) { return k == 0 || k == 3 || k == 6 ? 1 : (k == 2 || k == 5 ? 2 : (k == 7 ? 3 : 0)); }
__host__ __device__ constexpr int YF(int k) { return k == 1 || k == 3 || k == 5 ? 1 : (k == 4 || k == 6 ? 2 : (k == 8 ? 3 : 0)); }

// Geometry of an edgel's five points for the scatter: the edgel (m = 0) and
// its four NMS neighbours (m = 1..4), +-1 interp px along the gradient octant
// (cpu_toed.cpp:415-462): x-major iff |gx| >= |gy|; one step along +-gradient,
// axis and diagonal.  oct packs (x-major, gx < 0, gy < 0) in bits 0..2.
__device__ __forceinline__ void octant_neighbours(int oct, int nbr[4][2]) {
    const bool xm = oct & 1;
    const int sx = (oct & 2) ? -1 : 1, sy = (oct & 4) ? -1 : 1;
    nbr[0][0] = xm ? 0 : sy;   nbr[0][1] = xm ? sx : 0;
    nbr[1][0] = sy;            nbr[1][1] = sx;
    nbr[2][0] = xm ? 0 : -sy;  nbr[2][1] = xm ? -sx : 0;
    nbr[3][0] = -sy;           nbr[3][1] = -sx;
}

// Number of adjoint coefficients of an edgel: dL/d(nine responses at the
// edgel) (fx, fy, fxx, fxy, fyy, fxxy, fxyy, fxxx, fyyy), then dL/d(Ix, Iy)
// at the four neighbours.
constexpr int NADJ = 17;

// One edgel at interp pixel (R, C), upstream row `up` = dL/d[x, y,
// orientation, strength], lane l of a GL-lane group: the reverse pass from
// the edgel's outputs to the adjoints of the responses its outputs depend on.
// Tap tables, variants 0 unshifted, 1 unshifted with zero end taps (window R0
// of phase (0,0)), 2 shifted, each with one zero tap on either side so that
// index t + R1 + 1 is tap t = -R1-1 .. R1+1 (tap_tables_fill builds them):
// stp / sdp the FP64 taps and their sigma-derivatives, the four families (G
// to G''') of one tap together.  img(i, j) returns the pixel (0 outside the
// image).  Writes the NADJ coefficients (FP32, as the scatter uses them) to
// adjf and the octant to *oct on every lane of the group; lane 0 writes the
// edgel's dL/dsigma to *gs_out.  Returns false, writing nothing, for an edgel
// whose parabola is degenerate: it contributes no gradient.
//
// The edgel and its four neighbours (+-1 interp px = 1/2 image px) have their
// windows, and so their transposed filters, inside the P x P patch
// (P = 2 R1 + 2) whose first row is (R >> 1) - U + (R & 1) and first column
// (C >> 1) - U + (C & 1): over the patch every tap index is within
// -R1-1 .. R1+1.  Lane l of the group owns patch rows l + GL rr.
template <int R1, int GL = G, class Img>
__device__ __forceinline__ bool edgel_adjoint(int R, int C, float4 up, int l,
                                              const double (*stp)[2 * R1 + 3][4], const double (*sdp)[2 * R1 + 3][4],
                                              const Img &img, float *adjf, int *oct, double *gs_out) {
    constexpr int U = R1 + 1;
    constexpr int P = 2 * U;
    constexpr int NSR = (P + GL - 1) / GL;       // patch rows per lane
    const int pr0 = (R >> 1) - U + (R & 1), pc0 = (C >> 1) - U + (C & 1);

    // Each lane owns window rows p = -R1 + l + GL*rr (NRR of them); the column
    // loop is outermost so every tap read from LDS serves all of the lane's rows.
    constexpr int NRR = (2 * R1 + GL) / GL;

    // ---- 1a. nine responses at the edgel and their sigma-derivatives
    double v9[9], s9[9];
    {
        const int i = R >> 1, j = C >> 1, a = R & 1, b = C & 1, ph = 2 * a + b;
        const int vx = b ? 2 : (ph == 0 ? 1 : 0), vy = a ? 2 : (ph == 0 ? 1 : 0);
        double h[NRR][4], hs[NRR][4];
#pragma unroll
        for (int rr = 0; rr < NRR; ++rr)
#pragma unroll
            for (int f = 0; f < 4; ++f) { h[rr][f] = 0.0; hs[rr][f] = 0.0; }
#pragma unroll 3
        for (int q = -R1; q <= R1; ++q) {
            double xt[4], xs[4];
#pragma unroll
            for (int f = 0; f < 4; ++f) { xt[f] = stp[vx][q + R1 + 1][f]; xs[f] = sdp[vx][q + R1 + 1][f]; }
#pragma unroll
            for (int rr = 0; rr < NRR; ++rr) {
                const int p = -R1 + l + GL * rr;
                if (p > R1) continue;
                const double px = img(i - p, j - q);
#pragma unroll
                for (int f = 0; f < 4; ++f) {
                    h[rr][f] = fma(px, xt[f], h[rr][f]);
                    hs[rr][f] = fma(px, xs[f], hs[rr][f]);
                }
            }
        }
#pragma unroll
        for (int k = 0; k < 9; ++k) { v9[k] = 0.0; s9[k] = 0.0; }
#pragma unroll
        for (int rr = 0; rr < NRR; ++rr) {
            const int p = -R1 + l + GL * rr;
            if (p > R1) continue;
#pragma unroll
            for (int k = 0; k < 9; ++k) {
                const double y = stp[vy][p + R1 + 1][YF(k)], ys = sdp[vy][p + R1 + 1][YF(k)];
                v9[k] = fma(h[rr][XF(k)], y, v9[k]);
                s9[k] = fma(hs[rr][XF(k)], y, fma(h[rr][XF(k)], ys, s9[k]));
            }
        }
#pragma unroll
        for (int k = 0; k < 9; ++k) { v9[k] = toeddpp::group_sum<GL>(v9[k]); s9[k] = toeddpp::group_sum<GL>(s9[k]); }
    }
    // ---- 2a. reverse pass of the third-order orientation theta = atan2(TOx, -TOy)
    // (cpu_toed.cpp:227-232), which depends on the nine responses only:
    // its adjoints of the responses (ao, FP32 beyond fx and fy, which the
    // parabola also feeds) and their share of dL/dsigma
    double ao0 = 0.0, ao1 = 0.0, gs = 0.0;
    float ao[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) ao[k] = 0.f;
    {
        const double uth = up.z;
        const double fx = v9[0], fy = v9[1], fxx = v9[2], fxy = v9[3], fyy = v9[4];
        const double fxxy = v9[5], fxyy = v9[6], fxxx = v9[7], fyyy = v9[8];
        const double tox = fx * (2 * fxx * fxx + 2 * fxy * fxy) + fy * (2 * fxx * fxy + 2 * fyy * fxy)
                         + 2 * fx * fy * fxxy + fy * fy * fxyy + fx * fx * fxxx;
        const double toy = fx * (2 * fxx * fxy + 2 * fyy * fxy) + fy * (2 * fyy * fyy + 2 * fxy * fxy)
                         + 2 * fx * fy * fxyy + fx * fx * fxxy + fy * fy * fyyy;
        const double den = tox * tox + toy * toy;
        if (den > 0.0 && uth != 0.0) {
            // d atan2(u, v) = (v du - u dv) / (u^2 + v^2) with u = TOx, v = -TOy
            const double rden = uth / den;
            const double txb = -toy * rden, tyb = tox * rden;
            double a[9];
            a[0] = txb * (2 * fxx * fxx + 2 * fxy * fxy + 2 * fy * fxxy + 2 * fx * fxxx)
                 + tyb * (2 * fxx * fxy + 2 * fyy * fxy + 2 * fy * fxyy + 2 * fx * fxxy);
            a[1] = txb * (2 * fxx * fxy + 2 * fyy * fxy + 2 * fx * fxxy + 2 * fy * fxyy)
                 + tyb * (2 * fyy * fyy + 2 * fxy * fxy + 2 * fx * fxyy + 2 * fy * fyyy);
            a[2] = txb * (4 * fx * fxx + 2 * fy * fxy) + tyb * (2 * fx * fxy);
            a[3] = txb * (4 * fx * fxy + 2 * fy * fxx + 2 * fy * fyy) + tyb * (2 * fx * fxx + 2 * fx * fyy + 4 * fy * fxy);
            a[4] = txb * (2 * fy * fxy) + tyb * (2 * fx * fxy + 4 * fy * fyy);
            a[5] = txb * (2 * fx * fy) + tyb * (fx * fx);
            a[6] = txb * (fy * fy) + tyb * (2 * fx * fy);
            a[7] = txb * (fx * fx);
            a[8] = tyb * (fy * fy);
            ao0 = a[0];
            ao1 = a[1];
#pragma unroll
            for (int k = 2; k < 9; ++k) {
                gs = fma(a[k], s9[k], gs);
                ao[k] = (float)a[k];
            }
        }
    }

    // octant of the recomputed gradient (cpu_toed.cpp:415-462): x-major iff
    // |gx| >= |gy|; neighbours one step along +-gradient, axis and diagonal
    const double gx = v9[0], gy = v9[1];
    const int sx = gx >= 0.0 ? 1 : -1, sy = gy >= 0.0 ? 1 : -1;
    const bool xm = fabs(gx) >= fabs(gy);
    const int oc = (xm ? 1 : 0) | (sx < 0 ? 2 : 0) | (sy < 0 ? 4 : 0);
    int nbr[4][2];
    octant_neighbours(oc, nbr);

    // ---- 1b. Ix, Iy at the neighbours (+ sigma-derivatives): two passes over
    // the patch columns, two neighbours each (a lane row's image value serves
    // both), so that the row sums of only two neighbours are live at a time
    double nv[4][2], ns[4][2];
#pragma unroll
    for (int mp = 0; mp < 4; mp += 2) {
        int qb[2], vxn[2];
#pragma unroll
        for (int mm = 0; mm < 2; ++mm) {
            const int m = mp + mm;
            const int Cn = C + nbr[m][1], jn = Cn >> 1;
            vxn[mm] = (Cn & 1) ? 2 : (((R + nbr[m][0]) & 1) == 0 ? 1 : 0);
            qb[mm] = jn - pc0 + R1 + 1;          // - c: tap column index q = jn - image col
        }
        double h0[2][NSR], h1[2][NSR], hs0[2][NSR], hs1[2][NSR];
#pragma unroll
        for (int mm = 0; mm < 2; ++mm)
#pragma unroll
            for (int rr = 0; rr < NSR; ++rr) { h0[mm][rr] = 0; h1[mm][rr] = 0; hs0[mm][rr] = 0; hs1[mm][rr] = 0; }
#pragma unroll 4
        for (int c = 0; c < P; ++c) {
            double px[NSR];
#pragma unroll
            for (int rr = 0; rr < NSR; ++rr) {
                const int r = l + GL * rr;
                px[rr] = r < P ? img(pr0 + r, pc0 + c) : 0.0;
            }
#pragma unroll
            for (int mm = 0; mm < 2; ++mm) {
                const double *xt = stp[vxn[mm]][qb[mm] - c], *xs = sdp[vxn[mm]][qb[mm] - c];
                const double x0 = xt[0], x1 = xt[1], xs0 = xs[0], xs1 = xs[1];
#pragma unroll
                for (int rr = 0; rr < NSR; ++rr) {
                    h0[mm][rr] = fma(px[rr], x0, h0[mm][rr]);
                    h1[mm][rr] = fma(px[rr], x1, h1[mm][rr]);
                    hs0[mm][rr] = fma(px[rr], xs0, hs0[mm][rr]);
                    hs1[mm][rr] = fma(px[rr], xs1, hs1[mm][rr]);
                }
            }
        }
#pragma unroll
        for (int mm = 0; mm < 2; ++mm) {
            const int m = mp + mm;
            const int Rn = R + nbr[m][0], in = Rn >> 1;
            const int vy = (Rn & 1) ? 2 : (((C + nbr[m][1]) & 1) == 0 ? 1 : 0);
            double ix = 0, iy = 0, ixs = 0, iys = 0;
#pragma unroll
            for (int rr = 0; rr < NSR; ++rr) {
                const int r = l + GL * rr;
                if (r >= P) continue;
                const double *yt = stp[vy][in - (pr0 + r) + R1 + 1], *ys = sdp[vy][in - (pr0 + r) + R1 + 1];
                const double y0 = yt[0], y1 = yt[1], ys0 = ys[0], ys1 = ys[1];
                ix = fma(h1[mm][rr], y0, ix);
                iy = fma(h0[mm][rr], y1, iy);
                ixs = fma(hs1[mm][rr], y0, fma(h1[mm][rr], ys0, ixs));
                iys = fma(hs0[mm][rr], y1, fma(h0[mm][rr], ys1, iys));
            }
            nv[m][0] = toeddpp::group_sum<GL>(ix); nv[m][1] = toeddpp::group_sum<GL>(iy);
            ns[m][0] = toeddpp::group_sum<GL>(ixs); ns[m][1] = toeddpp::group_sum<GL>(iys);
        }
    }

    // ---- 2b. reverse pass of the parabola fit, subpixel position and strength
    // (every lane holds the totals; all compute it), reciprocals taken once
    const double ux = up.x, uy = up.y, us = up.w;
    const double f = sqrt(gx * gx + gy * gy), rf = 1.0 / f;
    const double nx = gx * rf, ny = gy * rf;
    const double s = xm ? (sy * gy) / (sx * gx) : (sx * gx) / (sy * gy);
    double M[4];
#pragma unroll
    for (int m = 0; m < 4; ++m) M[m] = sqrt(nv[m][0] * nv[m][0] + nv[m][1] * nv[m][1]);
    const double fp = M[0] + s * (M[1] - M[0]), fm = M[2] + s * (M[3] - M[2]);
    const double q = 1.0 + s * s, sq = sqrt(q), rq = 1.0 / q, rsq = 1.0 / sq;
    const double A = (fm + fp - 2.0 * f) * (0.5 * rq);
    if (!(isfinite(A) && A != 0.0 && f > 0.0)) return false;
    const double rA = 1.0 / A;
    const double B = (fp - fm) * (0.5 * rsq);
    const double t = -0.5 * B * rA;
    const double maxf = f - 0.25 * B * B * rA;

    const double tbar = 0.5 * (ux * nx + uy * ny);
    const double nxbar = 0.5 * ux * t, nybar = 0.5 * uy * t;
    const double mbar = us * (maxf >= 0.0 ? 1.0 : -1.0);
    const double Abar = (0.25 * mbar * B * B + 0.5 * tbar * B) * (rA * rA);
    const double Bbar = -(0.5 * mbar * B + 0.5 * tbar) * rA;
    double fbar = mbar - Abar * rq;
    const double fpbar = 0.5 * Abar * rq + 0.5 * Bbar * rsq;
    const double fmbar = 0.5 * Abar * rq - 0.5 * Bbar * rsq;
    const double sqbar = -Bbar * B * rsq;
    const double qbar = -Abar * A * rq + 0.5 * sqbar * rsq;
    const double sbar = 2.0 * s * qbar + fpbar * (M[1] - M[0]) + fmbar * (M[3] - M[2]);
    double gxbar = ao0, gybar = ao1;
    if (xm) { const double rg = 1.0 / gx; gybar += sbar * sy * sx * rg; gxbar -= sbar * s * rg; }
    else    { const double rg = 1.0 / gy; gxbar += sbar * sx * sy * rg; gybar -= sbar * s * rg; }
    gxbar += nxbar * rf;
    gybar += nybar * rf;
    fbar -= (nxbar * gx + nybar * gy) * (rf * rf);
    gxbar += fbar * gx * rf;
    gybar += fbar * gy * rf;
    // neighbour adjoints: M = |(Ix, Iy)|
    double an[8];
#pragma unroll
    for (int m = 0; m < 4; ++m) {
        const double Mb = (m == 0) ? fpbar * (1.0 - s) : (m == 1) ? fpbar * s : (m == 2) ? fmbar * (1.0 - s) : fmbar * s;
        const double w = M[m] > 0.0 ? Mb / M[m] : 0.0;
        an[2 * m] = w * nv[m][0];
        an[2 * m + 1] = w * nv[m][1];
    }

    // ---- 3. dL/dsigma (summed per workgroup by the caller)
    if (l == 0) {
        gs = fma(gxbar, s9[0], fma(gybar, s9[1], gs));
#pragma unroll
        for (int m = 0; m < 4; ++m) gs = fma(an[2 * m], ns[m][0], fma(an[2 * m + 1], ns[m][1], gs));
        *gs_out = gs;
    }

    adjf[0] = (float)gxbar;
    adjf[1] = (float)gybar;
#pragma unroll
    for (int k = 2; k < 9; ++k) adjf[k] = ao[k];
#pragma unroll
    for (int k = 9; k < NADJ; ++k) adjf[k] = (float)an[k - 9];
    *oct = oc;
    return true;
}

// The y-tap coefficients of point m of an edgel at one patch row, per x
// family: the point's y taps at the row (yv, families G .. G''') times its
// adjoints.  The edgel uses all four x families, a neighbour two (Ix is G'
// in x and G in y, Iy G in x and G' in y).
__device__ __forceinline__ void point_row_coefs(int m, const float *adjf, const float yv[4], float cf[4]) {
#pragma unroll
    for (int f = 0; f < 4; ++f) cf[f] = 0.f;
    if (m == 0) {
#pragma unroll
        for (int k = 0; k < 9; ++k) cf[XF(k)] = fmaf(adjf[k], yv[YF(k)], cf[XF(k)]);
    } else {
        cf[1] = adjf[7 + 2 * m] * yv[0];
        cf[0] = adjf[8 + 2 * m] * yv[1];
    }
}

// Tap-table indices of the five points relative to the edgel's P x P patch
// (first row pr0, first column pc0): variant and base of the x taps (tap
// column index q = jp - image col is stored at q + R1 + 1, so patch column c
// reads qb - c) and of the y taps (patch row r reads yb - r).
template <int R1>
__device__ __forceinline__ void point_taps(int R, int C, int oct, int vx[5], int qb[5], int vy[5], int yb[5]) {
    constexpr int U = R1 + 1;
    const int pr0 = (R >> 1) - U + (R & 1), pc0 = (C >> 1) - U + (C & 1);
    int nbr[4][2];
    octant_neighbours(oct, nbr);
#pragma unroll
    for (int m = 0; m < 5; ++m) {
        const int Rp = m == 0 ? R : R + nbr[m - 1][0], Cp = m == 0 ? C : C + nbr[m - 1][1];
        const int ip = Rp >> 1, jp = Cp >> 1, a = Rp & 1, b = Cp & 1, ph = 2 * a + b;
        vx[m] = b ? 2 : (ph == 0 ? 1 : 0);
        vy[m] = a ? 2 : (ph == 0 ? 1 : 0);
        qb[m] = jp - pc0 + R1 + 1;
        yb[m] = ip - pr0 + R1 + 1;
    }
}

// Scatter of one edgel's transposed filters into dL/dimage over its patch by
// its lane group: per point and patch column one float4 of x taps serves all
// of the lane's rows; each column's sums go to sink(i, j, v) (pixels of the
// H x W image only).  stf4: the FP32 copy of stp (the contributions are
// rounded to FP32 anyway).
template <int R1, int GL = G, class Sink>
__device__ __forceinline__ void edgel_scatter(int R, int C, int oct, const float *adjf, int l, int H, int W,
                                              const float4 (*stf4)[2 * R1 + 3], const Sink &sink) {
    constexpr int U = R1 + 1;
    constexpr int P = 2 * U;
    constexpr int NSR = (P + GL - 1) / GL;
    const int pr0 = (R >> 1) - U + (R & 1), pc0 = (C >> 1) - U + (C & 1);
    int vxm[5], qb[5], vym[5], yb[5];
    point_taps<R1>(R, C, oct, vxm, qb, vym, yb);
    float cf[5][NSR][4];
#pragma unroll
    for (int m = 0; m < 5; ++m) {
#pragma unroll
        for (int rr = 0; rr < NSR; ++rr) {
            const int r = l + GL * rr;
            const float4 y4 = r < P ? stf4[vym[m]][yb[m] - r] : make_float4(0.f, 0.f, 0.f, 0.f);
            const float yv[4] = {y4.x, y4.y, y4.z, y4.w};
            point_row_coefs(m, adjf, yv, cf[m][rr]);
        }
    }
#pragma unroll 2
    for (int c = 0; c < P; ++c) {
        float acc[NSR];
#pragma unroll
        for (int rr = 0; rr < NSR; ++rr) acc[rr] = 0.f;
#pragma unroll
        for (int m = 0; m < 5; ++m) {
            const float4 x4 = stf4[vxm[m]][qb[m] - c];
#pragma unroll
            for (int rr = 0; rr < NSR; ++rr) {
                if (m == 0)
                    acc[rr] = fmaf(cf[0][rr][0], x4.x, fmaf(cf[0][rr][1], x4.y, fmaf(cf[0][rr][2], x4.z, fmaf(cf[0][rr][3], x4.w, acc[rr]))));
                else
                    acc[rr] = fmaf(cf[m][rr][0], x4.x, fmaf(cf[m][rr][1], x4.y, acc[rr]));
            }
        }
        const int jj = pc0 + c;
        if (jj < 0 || jj >= W) continue;
#pragma unroll
        for (int rr = 0; rr < NSR; ++rr) {
            const int r = l + GL * rr, ii = pr0 + r;
            if (r < P && ii >= 0 && ii < H && acc[rr] != 0.f) sink(ii, jj, acc[rr]);
        }
    }
}

// The contribution to one patch pixel from its row's coefficients cf (from
// point_row_coefs, per point) and the x taps at its column c: the same FP32
// operations in the same order as edgel_scatter, for scatters that give a
// lane pixels of one edgel at a time.  vx, qb from point_taps.
template <int R1>
__device__ __forceinline__ float patch_column(const float (*cf)[4], const int vx[5], const int qb[5], int c,
                                              const float4 (*stf4)[2 * R1 + 3]) {
    float acc = 0.f;
#pragma unroll
    for (int m = 0; m < 5; ++m) {
        const float4 x4 = stf4[vx[m]][qb[m] - c];
        if (m == 0)
            acc = fmaf(cf[0][0], x4.x, fmaf(cf[0][1], x4.y, fmaf(cf[0][2], x4.z, fmaf(cf[0][3], x4.w, acc))));
        else
            acc = fmaf(cf[m][0], x4.x, fmaf(cf[m][1], x4.y, acc));
    }
    return acc;
}

// Adjoint and scatter of one edgel by its lane group; sink(i, j, v) adds v to
// dL/dimage.
template <int R1, int GL = G, class Img, class Sink>
__device__ __forceinline__ void edgel_backward(int R, int C, float4 up, int l, int H, int W,
                                               const double (*stp)[2 * R1 + 3][4], const double (*sdp)[2 * R1 + 3][4],
                                               const float4 (*stf4)[2 * R1 + 3], const Img &img, const Sink &sink,
                                               double *gs_out) {
    float adjf[NADJ];
    int oct;
    if (edgel_adjoint<R1, GL>(R, C, up, l, stp, sdp, img, adjf, &oct, gs_out))
        edgel_scatter<R1, GL>(R, C, oct, adjf, l, H, W, stf4, sink);
}

// The tap tables of edgel_backward from tap(v, f, t) and dtap(v, f, t)
// (variant v, family f, t = 0 .. 2 R1), threads tid of nt: zero border at
// index 0 and 2 R1 + 2.
template <int R1, class TapAt, class DTapAt>
__device__ __forceinline__ void tap_tables_fill(double (*stp)[2 * R1 + 3][4], double (*sdp)[2 * R1 + 3][4],
                                                float4 (*stf4)[2 * R1 + 3], int tid, int nt, const TapAt &tap,
                                                const DTapAt &dtap) {
    constexpr int NP = 2 * R1 + 3;
    for (int k = tid; k < 3 * NP; k += nt) {
        const int v = k / NP, i = k - v * NP, t = i - 1;
        const bool in = t >= 0 && t <= 2 * R1;
        double a[4];
#pragma unroll
        for (int f = 0; f < 4; ++f) {
            a[f] = in ? tap(v, f, t) : 0.0;
            stp[v][i][f] = a[f];
            sdp[v][i][f] = in ? dtap(v, f, t) : 0.0;
        }
        stf4[v][i] = make_float4((float)a[0], (float)a[1], (float)a[2], (float)a[3]);
    }
}

}  // namespace toedadj
