// Per-edgel reverse pass of the TOED outputs with the edgel set frozen, shared
// by the index-driven backward (toed_backward.hip) and the tile-driven
// backward of a batch plan (toed_detect.hip).  FP64 throughout; one G-lane
// group per edgel (lanes split the window rows, totals by __shfl_xor).
//
// Differentiated forward arithmetic: the nine responses at a pixel
// (cpu_toed.cpp:194-221), octant choice and neighbour interpolation
// (:415-462), parabola and subpixel position (:465-489), strength (:481-485),
// third-order orientation (:227-232).
#pragma once

#include <hip/hip_runtime.h>

namespace toedadj {

constexpr int G = 8;    // lanes per edgel (window rows split over the group)

__device__ __forceinline__ double redG(double v) {
#pragma unroll
    for (int m = G / 2; m >= 1; m >>= 1) v += __shfl_xor(v, m, G);
    return v;
}

// response k (fx, fy, fxx, fxy, fyy, fxxy, fxyy, fxxx, fyyy) -> x (column) and
// y (row) filter family (0 G, 1 G', 2 G'', 3 G''')
__host__ __device__ constexpr int XF(int k) { return k == 0 || k == 3 || k == 6 ? 1 : (k == 2 || k == 5 ? 2 : (k == 7 ? 3 : 0)); }
__host__ __device__ constexpr int YF(int k) { return k == 1 || k == 3 || k == 5 ? 1 : (k == 4 || k == 6 ? 2 : (k == 8 ? 3 : 0)); }

// One edgel at interp pixel (R, C) of an H x W image, upstream row `up` =
// dL/d[x, y, orientation, strength], group lane l.  st / sd: FP64 taps and
// their sigma-derivatives, variants 0 unshifted, 1 unshifted with zero end
// taps (window R0 of phase (0,0)), 2 shifted; stf the FP32 copy of st used by
// the scatter (its contributions are rounded to FP32 for the atomics anyway).
// img(i, j) returns the pixel
// (0 outside the image); sink(i, j, v) adds v to dL/dimage.  Lane 0 writes
// the edgel's dL/dsigma to *gs_out.
template <int R1, class Img, class Sink>
__device__ __forceinline__ void edgel_backward(int R, int C, float4 up, int l, int H, int W,
                                               const double (*st)[4][2 * R1 + 1], const double (*sd)[4][2 * R1 + 1],
                                               const float (*stf)[4][2 * R1 + 1], const Img &img, const Sink &sink,
                                               double *gs_out) {
    constexpr int U = R1 + 1;          // patch half-width around the edgel (neighbours are +-1 interp px)

    // Each lane owns window rows p = -R1 + l + G*rr (NRR of them); the column
    // loop is outermost so every tap read from LDS serves all of the lane's rows.
    constexpr int NRR = (2 * R1 + G) / G;

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
#pragma unroll 1
        for (int q = -R1; q <= R1; ++q) {
            double xt[4], xs[4];
#pragma unroll
            for (int f = 0; f < 4; ++f) { xt[f] = st[vx][f][q + R1]; xs[f] = sd[vx][f][q + R1]; }
#pragma unroll
            for (int rr = 0; rr < NRR; ++rr) {
                const int p = -R1 + l + G * rr;
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
            const int p = -R1 + l + G * rr;
            if (p > R1) continue;
#pragma unroll
            for (int k = 0; k < 9; ++k) {
                const double y = st[vy][YF(k)][p + R1], ys = sd[vy][YF(k)][p + R1];
                v9[k] = fma(h[rr][XF(k)], y, v9[k]);
                s9[k] = fma(hs[rr][XF(k)], y, fma(h[rr][XF(k)], ys, s9[k]));
            }
        }
#pragma unroll
        for (int k = 0; k < 9; ++k) { v9[k] = redG(v9[k]); s9[k] = redG(s9[k]); }
    }
    // octant of the recomputed gradient (cpu_toed.cpp:415-462): x-major iff
    // |gx| >= |gy|; neighbours one step along +-gradient, axis and diagonal
    const double gx = v9[0], gy = v9[1];
    const int sx = gx >= 0.0 ? 1 : -1, sy = gy >= 0.0 ? 1 : -1;
    const bool xm = fabs(gx) >= fabs(gy);
    const int nbr[4][2] = {{xm ? 0 : sy, xm ? sx : 0}, {sy, sx}, {xm ? 0 : -sy, xm ? -sx : 0}, {-sy, -sx}};

    // ---- 1b. Ix, Iy at the neighbours (+ sigma-derivatives)
    double nv[4][2], ns[4][2];
#pragma unroll
    for (int m = 0; m < 4; ++m) {
        const int Rn = R + nbr[m][0], Cn = C + nbr[m][1];
        const int i = Rn >> 1, j = Cn >> 1, a = Rn & 1, b = Cn & 1, ph = 2 * a + b;
        const int vx = b ? 2 : (ph == 0 ? 1 : 0), vy = a ? 2 : (ph == 0 ? 1 : 0);
        double h0[NRR], h1[NRR], hs0[NRR], hs1[NRR];
#pragma unroll
        for (int rr = 0; rr < NRR; ++rr) { h0[rr] = 0; h1[rr] = 0; hs0[rr] = 0; hs1[rr] = 0; }
#pragma unroll 1
        for (int q = -R1; q <= R1; ++q) {
            const double x0 = st[vx][0][q + R1], x1 = st[vx][1][q + R1];
            const double xs0 = sd[vx][0][q + R1], xs1 = sd[vx][1][q + R1];
#pragma unroll
            for (int rr = 0; rr < NRR; ++rr) {
                const int p = -R1 + l + G * rr;
                if (p > R1) continue;
                const double px = img(i - p, j - q);
                h0[rr] = fma(px, x0, h0[rr]);
                h1[rr] = fma(px, x1, h1[rr]);
                hs0[rr] = fma(px, xs0, hs0[rr]);
                hs1[rr] = fma(px, xs1, hs1[rr]);
            }
        }
        double ix = 0, iy = 0, ixs = 0, iys = 0;
#pragma unroll
        for (int rr = 0; rr < NRR; ++rr) {
            const int p = -R1 + l + G * rr;
            if (p > R1) continue;
            const double y0 = st[vy][0][p + R1], y1 = st[vy][1][p + R1];
            const double ys0 = sd[vy][0][p + R1], ys1 = sd[vy][1][p + R1];
            ix = fma(h1[rr], y0, ix);
            iy = fma(h0[rr], y1, iy);
            ixs = fma(hs1[rr], y0, fma(h1[rr], ys0, ixs));
            iys = fma(hs0[rr], y1, fma(h0[rr], ys1, iys));
        }
        nv[m][0] = redG(ix); nv[m][1] = redG(iy);
        ns[m][0] = redG(ixs); ns[m][1] = redG(iys);
    }

    // ---- 2. reverse pass (every lane holds the totals; all compute it)
    const double ux = up.x, uy = up.y, uth = up.z, us = up.w;
    const double f = sqrt(gx * gx + gy * gy);
    const double nx = gx / f, ny = gy / f;
    const double s = xm ? (sy * gy) / (sx * gx) : (sx * gx) / (sy * gy);
    double M[4];
#pragma unroll
    for (int m = 0; m < 4; ++m) M[m] = sqrt(nv[m][0] * nv[m][0] + nv[m][1] * nv[m][1]);
    const double fp = M[0] + s * (M[1] - M[0]), fm = M[2] + s * (M[3] - M[2]);
    const double q = 1.0 + s * s, sq = sqrt(q);
    const double A = (fm + fp - 2.0 * f) / (2.0 * q);
    const double B = (fp - fm) / (2.0 * sq);
    const double t = -B / (2.0 * A);
    const double maxf = f - B * B / (4.0 * A);

    const double tbar = 0.5 * (ux * nx + uy * ny);
    double nxbar = 0.5 * ux * t, nybar = 0.5 * uy * t;
    const double mbar = us * (maxf >= 0.0 ? 1.0 : -1.0);
    const double Abar = mbar * B * B / (4.0 * A * A) + tbar * B / (2.0 * A * A);
    const double Bbar = -mbar * B / (2.0 * A) - tbar / (2.0 * A);
    double fbar = mbar - Abar / q;
    const double fpbar = Abar / (2.0 * q) + Bbar / (2.0 * sq);
    const double fmbar = Abar / (2.0 * q) - Bbar / (2.0 * sq);
    const double sqbar = -Bbar * B / sq;
    const double qbar = -Abar * A / q + sqbar / (2.0 * sq);
    double sbar = 2.0 * s * qbar + fpbar * (M[1] - M[0]) + fmbar * (M[3] - M[2]);
    const double Mbar[4] = {fpbar * (1.0 - s), fpbar * s, fmbar * (1.0 - s), fmbar * s};
    double gxbar = 0.0, gybar = 0.0;
    if (xm) { gybar += sbar * sy / (sx * gx); gxbar -= sbar * s / gx; }
    else    { gxbar += sbar * sx / (sy * gy); gybar -= sbar * s / gy; }
    gxbar += nxbar / f;
    gybar += nybar / f;
    fbar -= (nxbar * gx + nybar * gy) / (f * f);
    gxbar += fbar * gx / f;
    gybar += fbar * gy / f;
    double adj[17];
    adj[0] = gxbar; adj[1] = gybar;
#pragma unroll
    for (int k = 2; k < 9; ++k) adj[k] = 0.0;
#pragma unroll
    for (int m = 0; m < 4; ++m) {
        const double w = M[m] > 0.0 ? Mbar[m] / M[m] : 0.0;
        adj[9 + 2 * m] = w * nv[m][0];
        adj[10 + 2 * m] = w * nv[m][1];
    }
    {   // third-order orientation theta = atan2(TOx, -TOy) (cpu_toed.cpp:227-232)
        const double fx = v9[0], fy = v9[1], fxx = v9[2], fxy = v9[3], fyy = v9[4];
        const double fxxy = v9[5], fxyy = v9[6], fxxx = v9[7], fyyy = v9[8];
        const double tox = fx * (2 * fxx * fxx + 2 * fxy * fxy) + fy * (2 * fxx * fxy + 2 * fyy * fxy)
                         + 2 * fx * fy * fxxy + fy * fy * fxyy + fx * fx * fxxx;
        const double toy = fx * (2 * fxx * fxy + 2 * fyy * fxy) + fy * (2 * fyy * fyy + 2 * fxy * fxy)
                         + 2 * fx * fy * fxyy + fx * fx * fxxy + fy * fy * fyyy;
        const double den = tox * tox + toy * toy;
        if (den > 0.0 && uth != 0.0) {
            // d atan2(u, v) = (v du - u dv) / (u^2 + v^2) with u = TOx, v = -TOy
            const double txb = uth * (-toy) / den, tyb = uth * tox / den;
            adj[0] += txb * (2 * fxx * fxx + 2 * fxy * fxy + 2 * fy * fxxy + 2 * fx * fxxx)
                    + tyb * (2 * fxx * fxy + 2 * fyy * fxy + 2 * fy * fxyy + 2 * fx * fxxy);
            adj[1] += txb * (2 * fxx * fxy + 2 * fyy * fxy + 2 * fx * fxxy + 2 * fy * fxyy)
                    + tyb * (2 * fyy * fyy + 2 * fxy * fxy + 2 * fx * fxyy + 2 * fy * fyyy);
            adj[2] = txb * (4 * fx * fxx + 2 * fy * fxy) + tyb * (2 * fx * fxy);
            adj[3] = txb * (4 * fx * fxy + 2 * fy * fxx + 2 * fy * fyy) + tyb * (2 * fx * fxx + 2 * fx * fyy + 4 * fy * fxy);
            adj[4] = txb * (2 * fy * fxy) + tyb * (2 * fx * fxy + 4 * fy * fyy);
            adj[5] = txb * (2 * fx * fy) + tyb * (fx * fx);
            adj[6] = txb * (fy * fy) + tyb * (2 * fx * fy);
            adj[7] = txb * (fx * fx);
            adj[8] = tyb * (fy * fy);
        }
    }
    if (!(isfinite(A) && A != 0.0 && f > 0.0)) return;

    // ---- 3. dL/dsigma (summed per workgroup by the caller)
    if (l == 0) {
        double gs = 0.0;
#pragma unroll
        for (int k = 0; k < 9; ++k) gs = fma(adj[k], s9[k], gs);
#pragma unroll
        for (int m = 0; m < 4; ++m) gs = fma(adj[9 + 2 * m], ns[m][0], fma(adj[10 + 2 * m], ns[m][1], gs));
        *gs_out = gs;
    }

    // ---- 4. scatter the transposed filters into dL/dimage: lane rows of the
    // (2U+1)^2 patch around the edgel; per point and patch column the four x
    // taps are read once and applied to all of the lane's rows
    constexpr int NSR = (2 * U + G) / G;
    const int ic = R >> 1, jc = C >> 1;
    float adjf[17];
#pragma unroll
    for (int k = 0; k < 17; ++k) adjf[k] = (float)adj[k];
    float row[NSR][2 * U + 1];
#pragma unroll
    for (int rr = 0; rr < NSR; ++rr)
#pragma unroll
        for (int c = 0; c <= 2 * U; ++c) row[rr][c] = 0.f;
    // point m = -1: the edgel (nine responses); m >= 0: neighbour m (Ix, Iy)
#pragma unroll
    for (int m = -1; m < 4; ++m) {
        const int Rp = m < 0 ? R : R + nbr[m][0], Cp = m < 0 ? C : C + nbr[m][1];
        const int ip = Rp >> 1, jp = Cp >> 1, a = Rp & 1, b = Cp & 1, ph = 2 * a + b;
        const int vx = b ? 2 : (ph == 0 ? 1 : 0), vy = a ? 2 : (ph == 0 ? 1 : 0);
        float cf[NSR][4];                           // per lane row, per x-family coefficient
#pragma unroll
        for (int rr = 0; rr < NSR; ++rr) {
            const int p = ip - (ic - U + l + G * rr);   // tap row index: image row = ip - p
#pragma unroll
            for (int f = 0; f < 4; ++f) cf[rr][f] = 0.f;
            if (p < -R1 || p > R1) continue;
            if (m < 0) {
#pragma unroll
                for (int k = 0; k < 9; ++k) cf[rr][XF(k)] = fmaf(adjf[k], stf[vy][YF(k)][p + R1], cf[rr][XF(k)]);
            } else {
                cf[rr][1] = adjf[9 + 2 * m] * stf[vy][0][p + R1];     // Ix: G' in x, G in y
                cf[rr][0] = adjf[10 + 2 * m] * stf[vy][1][p + R1];    // Iy: G in x, G' in y
            }
        }
#pragma unroll
        for (int c = 0; c <= 2 * U; ++c) {
            const int qq = jp - (jc - U + c);          // tap column index: image col = jp - q
            if (qq < -R1 || qq > R1) continue;
            const float x0 = stf[vx][0][qq + R1], x1 = stf[vx][1][qq + R1];
            const float x2 = stf[vx][2][qq + R1], x3 = stf[vx][3][qq + R1];
#pragma unroll
            for (int rr = 0; rr < NSR; ++rr)
                row[rr][c] = fmaf(cf[rr][0], x0, fmaf(cf[rr][1], x1, fmaf(cf[rr][2], x2, fmaf(cf[rr][3], x3, row[rr][c]))));
        }
    }
#pragma unroll
    for (int rr = 0; rr < NSR; ++rr) {
        const int o = -U + l + G * rr;
        const int ii = ic + o;
        if (o > U || ii < 0 || ii >= H) continue;
#pragma unroll
        for (int c = 0; c <= 2 * U; ++c) {
            const int jj = jc - U + c;
            if (jj >= 0 && jj < W && row[rr][c] != 0.f) sink(ii, jj, row[rr][c]);
        }
    }
}

}  // namespace toedadj
