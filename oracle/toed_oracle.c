/*
 * TEST INFRASTRUCTURE ONLY -- the FP64 oracle for the TOED detection path.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load this code, and only as the checker (or the timed CPU baseline); the
 * product path never calls it.
 *
 * It restates the reference's double-precision CPU detector
 * (cpu_toed.cpp, class ThirdOrderEdgeDetectionCPU<double>) in plain C:
 *
 *   toed_oracle_taps      1-D Gaussian-derivative taps, cpu_toed.cpp:141-167
 *                         (closed forms of the commented-out generator at
 *                         :141-162; window radii R0 = ceil(4 sigma) for the
 *                         unshifted phase-0 window, R1 = R0 + 1, :131-132).
 *   toed_oracle_detect    convolve_img (cpu_toed.cpp:128-369), direct 2-D
 *                         sums per pixel and phase with the reference's tap
 *                         indexing, product grouping and summation order;
 *                         then non_maximum_suppresion (cpu_toed.cpp:379-506)
 *                         and the raster-order edgel list (:508-538).
 *
 * Pinned by: golden vectors of the reference compiled from its own sources
 * (oracle/_ref, tests/golden), closed-form tap values, and the analytic
 * checks in tests/test_oracle.py (impulse response, linear ramp, transpose
 * symmetry, translation, step-edge subpixel location, disk orientation).
 *
 * The reference hard-codes the sigma = 2 taps as 15-significant-digit
 * literals; this restatement evaluates the same closed forms in full double
 * precision, so against the compiled reference it agrees to ~1e-13 relative,
 * not bit for bit.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

#define ORACLE_MAX_R1 32

/* Window radii for sigma: R0 = ceil(4 sigma) (MATLAB Gx_2d_op.m:11 and the
 * 17-tap window of cpu_toed.cpp:205 at sigma = 2), R1 = R0 + 1
 * (cpu_toed.cpp:131-132). */
int toed_oracle_radius(double sigma, int *R0, int *R1) {
    if (!(sigma > 0.0)) return -1;
    int r0 = (int)ceil(4.0 * sigma - 1e-12);
    if (r0 < 1 || r0 + 1 > ORACLE_MAX_R1) return -1;
    *R0 = r0;
    *R1 = r0 + 1;
    return 0;
}

/* taps[f * n + k], n = 2*R1+1, k = 0..n-1, sample point x = k - R1 (+0.5 for
 * the shifted families).  f: 0 G, 1 G', 2 G'', 3 G''' (unshifted),
 * 4..7 the same at x + 1/2.  Closed forms follow cpu_toed.cpp:141-162. */
int toed_oracle_taps(double sigma, int *R0, int *R1, double *taps) {
    int r0, r1;
    if (toed_oracle_radius(sigma, &r0, &r1)) return -1;
    const int n = 2 * r1 + 1;
    const double s2 = sigma * sigma;
    const double norm = sqrt(2.0 * 3.14159265358979323846) * sigma;
    for (int shifted = 0; shifted < 2; ++shifted) {
        for (int k = 0; k < n; ++k) {
            double x = (double)(k - r1) + (shifted ? 0.5 : 0.0);
            double e = exp(-x * x / (2.0 * s2));
            double g = e / norm;
            taps[(4 * shifted + 0) * n + k] = g;
            taps[(4 * shifted + 1) * n + k] = -x * g / s2;
            taps[(4 * shifted + 2) * n + k] = (x * x - s2) * g / (s2 * s2);
            taps[(4 * shifted + 3) * n + k] = x * (3.0 * s2 - x * x) * g / (s2 * s2 * s2);
        }
    }
    if (R0) *R0 = r0;
    if (R1) *R1 = r1;
    return 0;
}

/* The nine responses at one interp sample, cpu_toed.cpp:194-221 etc. */
typedef struct { double fx, fy, fxx, fxy, fyy, fxxy, fxyy, fxxx, fyyy; } resp_t;

/*
 * One phase of cpu_toed.cpp's convolution at original pixel (i, j).
 *   ph = 0: 17x17 window (radius R0), unshifted taps, tap index p + R0 + 1
 *           (cpu_toed.cpp:205-221); fx/fy multiply img by the tap product.
 *   ph = 1: x shifted by +1/2 (cpu_toed.cpp:246-262)
 *   ph = 2: y shifted by +1/2 (cpu_toed.cpp:287-303)
 *   ph = 3: both shifted    (cpu_toed.cpp:328-344)
 *   phases 1-3: radius R1, tap index p + R1, (img * A) * B grouping.
 * q indexes the x (column) taps, p the y (row) taps, p outer, q inner.
 */
static void conv_phase(const double *img, int H, int W, int i, int j, int ph,
                       int R0, int R1, const double *tp, resp_t *o) {
    const int n = 2 * R1 + 1;
    const double *ux = tp + ((ph & 1) ? 4 : 0) * n;   /* x (q) families */
    const double *uy = tp + ((ph & 2) ? 4 : 0) * n;   /* y (p) families */
    const double *Gq = ux, *G1q = ux + n, *G2q = ux + 2 * n, *G3q = ux + 3 * n;
    const double *Gp = uy, *G1p = uy + n, *G2p = uy + 2 * n, *G3p = uy + 3 * n;
    const int rad = (ph == 0) ? R0 : R1;
    double fx = 0, fy = 0, fxx = 0, fxy = 0, fyy = 0, fxxy = 0, fxyy = 0, fxxx = 0, fyyy = 0;
    for (int p = -rad; p <= rad; ++p) {
        const int ii = i - p;
        if (ii < 0 || ii >= H) continue;
        const int kp = p + R1;   /* == p + R0 + 1 for phase 0 */
        for (int q = -rad; q <= rad; ++q) {
            const int jj = j - q;
            if (jj < 0 || jj >= W) continue;
            const int kq = q + R1;
            const double v = img[(size_t)ii * W + jj];
            if (ph == 0) {
                fx += v * (G1q[kq] * Gp[kp]);
                fy += v * (Gq[kq] * G1p[kp]);
            } else {
                fx += (v * G1q[kq]) * Gp[kp];
                fy += (v * Gq[kq]) * G1p[kp];
            }
            fxx  += (v * G2q[kq]) * Gp[kp];
            fxy  += (v * G1q[kq]) * G1p[kp];
            fyy  += (v * Gq[kq]) * G2p[kp];
            fxxy += (v * G2q[kq]) * G1p[kp];
            fxyy += (v * G1q[kq]) * G2p[kp];
            fxxx += (v * G3q[kq]) * Gp[kp];
            fyyy += (v * Gq[kq]) * G3p[kp];
        }
    }
    o->fx = fx; o->fy = fy; o->fxx = fxx; o->fxy = fxy; o->fyy = fyy;
    o->fxxy = fxxy; o->fxyy = fxyy; o->fxxx = fxxx; o->fyyy = fyyy;
}

/* Third-order orientation, cpu_toed.cpp:227-232: theta = atan2 of the
 * normalised gradient of (grad I)^T H (grad I), same evaluation order. */
double toed_oracle_to_theta(double fx, double fy, double fxx, double fxy, double fyy,
                            double fxxy, double fxyy, double fxxx, double fyyy) {
    double tx = fx * (2 * fxx * fxx + 2 * fxy * fxy) + fy * (2 * fxx * fxy + 2 * fyy * fxy)
              + 2 * fx * fy * fxxy + fy * fy * fxyy + fx * fx * fxxx;
    double ty = fx * (2 * fxx * fxy + 2 * fyy * fxy) + fy * (2 * fyy * fyy + 2 * fxy * fxy)
              + 2 * fx * fy * fxyy + fx * fx * fxxy + fy * fy * fyyy;
    double m = sqrt(tx * tx + ty * ty);
    tx /= m;
    ty /= m;
    return atan2(tx, -ty);
}

/*
 * Convolution stage (cpu_toed.cpp:128-369): fills the four (2H)x(2W) planes.
 * Optional resp9 receives all nine responses per interp pixel
 * (9 x (2H)(2W), derivative-major) for tests.
 */
int toed_oracle_convolve(const double *img, int H, int W, double sigma, int nthreads,
                         double *Ix, double *Iy, double *mag, double *orient, double *resp9) {
    int R0, R1;
    double tp[8 * (2 * ORACLE_MAX_R1 + 1)];
    if (H <= 0 || W <= 0 || !img || !Ix || !Iy || !mag || !orient) return -1;
    if (toed_oracle_taps(sigma, &R0, &R1, tp)) return -1;
    const int IW = 2 * W;
    const size_t plane = (size_t)(2 * H) * IW;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
#pragma omp parallel for schedule(dynamic)
    for (int i = 0; i < H; ++i) {
        for (int j = 0; j < W; ++j) {
            for (int ph = 0; ph < 4; ++ph) {
                resp_t r;
                conv_phase(img, H, W, i, j, ph, R0, R1, tp, &r);
                const size_t o = (size_t)(2 * i + (ph >> 1)) * IW + (2 * j + (ph & 1));
                Ix[o] = r.fx;
                Iy[o] = r.fy;
                mag[o] = sqrt(r.fx * r.fx + r.fy * r.fy);
                orient[o] = toed_oracle_to_theta(r.fx, r.fy, r.fxx, r.fxy, r.fyy,
                                                 r.fxxy, r.fxyy, r.fxxx, r.fyyy);
                if (resp9) {
                    const double v[9] = {r.fx, r.fy, r.fxx, r.fxy, r.fyy, r.fxxy, r.fxyy, r.fxxx, r.fyyy};
                    for (int d = 0; d < 9; ++d) resp9[d * plane + o] = v[d];
                }
            }
        }
    }
    return 0;
}

/*
 * Octant selection and neighbour interpolation of cpu_toed.cpp:415-462 at
 * interp pixel (i, j).  Returns 0 when no branch applies (NaN gradient; the
 * reference then reads uninitialised values, here the pixel is rejected).
 * Writes slope, fp, fm and the octant (1..8) in the reference's numbering.
 */
static int octant_interp(const double *Ix, const double *Iy, const double *M, int IW,
                         int i, int j, double nx, double ny,
                         double *slope, double *fp, double *fm, int *oct) {
#define MG(a, b) M[(size_t)(a) * IW + (b)]
    const double gx = Ix[(size_t)i * IW + j], gy = Iy[(size_t)i * IW + j];
    double s;
    if (gx >= 0 && gy >= 0) {
        if (gx >= gy) {
            s = ny / nx; *oct = 1;
            *fp = MG(i, j + 1) * (1 - s) + MG(i + 1, j + 1) * s;
            *fm = MG(i, j - 1) * (1 - s) + MG(i - 1, j - 1) * s;
        } else {
            s = nx / ny; *oct = 2;
            *fp = MG(i + 1, j) * (1 - s) + MG(i + 1, j + 1) * s;
            *fm = MG(i - 1, j) * (1 - s) + MG(i - 1, j - 1) * s;
        }
    } else if (gx < 0 && gy >= 0) {
        if (fabs(gx) < gy) {
            s = -nx / ny; *oct = 3;
            *fp = MG(i + 1, j) * (1 - s) + MG(i + 1, j - 1) * s;
            *fm = MG(i - 1, j) * (1 - s) + MG(i - 1, j + 1) * s;
        } else {
            s = -ny / nx; *oct = 4;
            *fp = MG(i, j - 1) * (1 - s) + MG(i + 1, j - 1) * s;
            *fm = MG(i, j + 1) * (1 - s) + MG(i - 1, j + 1) * s;
        }
    } else if (gx < 0 && gy < 0) {
        if (fabs(gx) >= fabs(gy)) {
            s = ny / nx; *oct = 5;
            *fp = MG(i, j - 1) * (1 - s) + MG(i - 1, j - 1) * s;
            *fm = MG(i, j + 1) * (1 - s) + MG(i + 1, j + 1) * s;
        } else {
            s = nx / ny; *oct = 6;
            *fp = MG(i - 1, j) * (1 - s) + MG(i - 1, j - 1) * s;
            *fm = MG(i + 1, j) * (1 - s) + MG(i + 1, j + 1) * s;
        }
    } else if (gx >= 0 && gy < 0) {
        if (gx < fabs(gy)) {
            s = -nx / ny; *oct = 7;
            *fp = MG(i - 1, j) * (1 - s) + MG(i - 1, j + 1) * s;
            *fm = MG(i + 1, j) * (1 - s) + MG(i + 1, j - 1) * s;
        } else {
            s = -ny / nx; *oct = 8;
            *fp = MG(i, j + 1) * (1 - s) + MG(i - 1, j + 1) * s;
            *fm = MG(i, j - 1) * (1 - s) + MG(i + 1, j - 1) * s;
        }
    } else {
        return 0;
    }
#undef MG
    *slope = s;
    return 1;
}

/*
 * NMS + list (cpu_toed.cpp:379-545) over precomputed planes.
 *   edges   N x 4 [x, y, theta, strength] in original-pixel units
 *           (x = (x_interp - 1)/2, cpu_toed.cpp:516-529); rows beyond cap are
 *           counted but not written.
 *   eidx    optional, interp linear index i*(2W)+j of each edgel.
 *   mmag    optional (2H)(2W) plane: the smallest |difference| among the
 *           magnitude comparisons decided at the pixel (threshold, peak
 *           test); +inf where none is evaluated.
 *   msub    optional plane: | |s*| - sqrt(2) | where the peak test passed.
 * Returns the edgel count.
 */
long toed_oracle_nms(const double *Ix, const double *Iy, const double *mag, const double *orient,
                     int H, int W, double thresh, int border, int nthreads,
                     double *edges, long cap, int64_t *eidx, double *mmag, double *msub) {
    const int IH = 2 * H, IW = 2 * W;
    const size_t plane = (size_t)IH * IW;
    double *px = (double *)calloc(plane, sizeof(double));
    double *py = (double *)calloc(plane, sizeof(double));
    double *pm = (double *)calloc(plane, sizeof(double));
    if (!px || !py || !pm) { free(px); free(py); free(pm); return -1; }
    if (mmag) for (size_t k = 0; k < plane; ++k) mmag[k] = INFINITY;
    if (msub) for (size_t k = 0; k < plane; ++k) msub[k] = INFINITY;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
#pragma omp parallel for schedule(dynamic)
    for (int j = border; j < IW - border; ++j) {
        for (int i = border; i < IH - border; ++i) {
            const size_t o = (size_t)i * IW + j;
            const double f = mag[o];
            if (mmag) mmag[o] = fabs(f - thresh);
            if (f <= thresh) continue;                                   /* :403 */
            if (fabs(Ix[o]) < 10e-6 && fabs(Iy[o]) < 10e-6) continue;    /* :407 */
            const double nx = Ix[o] / f, ny = Iy[o] / f;                 /* :411-412 */
            double slope, fp, fm; int oct;
            if (!octant_interp(Ix, Iy, mag, IW, i, j, nx, ny, &slope, &fp, &fm, &oct)) continue;
            if (mmag) {
                double m = fabs(f - fm);
                if (fabs(f - fp) < m) m = fabs(f - fp);
                if (mmag[o] < m) m = mmag[o];
                mmag[o] = m;
            }
            const double s = sqrt(1 + slope * slope);                    /* :465 */
            if ((f > fm && f > fp) || (f > fm && f >= fp) || (f >= fm && f > fp)) {
                const double A = (fm + fp - 2 * f) / (2 * s * s);        /* :471-476 */
                const double B = (fp - fm) / (2 * s);
                const double C = f;
                const double ss = -B / (2 * A);
                const double maxf = A * ss * ss + B * ss + C;
                if (msub) msub[o] = fabs(fabs(ss) - sqrt(2.0));
                if (fabs(ss) <= sqrt(2.0)) {                             /* :478 */
                    const double gxs = maxf * nx, gys = maxf * ny;
                    pm[o] = sqrt(gxs * gxs + gys * gys);                 /* :481-485 */
                    px[o] = j + ss * nx;                                 /* :488-489 */
                    py[o] = i + ss * ny;
                }
            }
        }
    }
    long n = 0;
    for (int i = 0; i < IH; ++i) {                                       /* :510-538 */
        for (int j = 0; j < IW; ++j) {
            const size_t o = (size_t)i * IW + j;
            if (px[o] != 0) {
                if (n < cap && edges) {
                    edges[4 * n + 0] = (px[o] - 1) / 2;
                    edges[4 * n + 1] = (py[o] - 1) / 2;
                    edges[4 * n + 2] = orient[o];
                    edges[4 * n + 3] = pm[o];
                }
                if (n < cap && eidx) eidx[n] = (int64_t)o;
                ++n;
            }
        }
    }
    free(px); free(py); free(pm);
    return n;
}

/* Whole detector: convolve_img + non_maximum_suppresion with the reference's
 * constants (threshold 2 at cpu_toed.cpp:403, border 10 at :400-401).
 * Planes may be NULL (allocated internally). Returns N or -1. */
long toed_oracle_detect(const double *img, int H, int W, double sigma, int nthreads,
                        double thresh, int border,
                        double *edges, long cap, int64_t *eidx,
                        double *Ix, double *Iy, double *mag, double *orient,
                        double *mmag, double *msub) {
    if (H <= 0 || W <= 0 || !img) return -1;
    const size_t plane = (size_t)(2 * H) * (2 * W);
    double *bufs[4] = {Ix, Iy, mag, orient};
    double *own[4] = {0, 0, 0, 0};
    for (int k = 0; k < 4; ++k) {
        if (!bufs[k]) {
            own[k] = (double *)malloc(plane * sizeof(double));
            if (!own[k]) { for (int q = 0; q < 4; ++q) free(own[q]); return -1; }
            bufs[k] = own[k];
        }
    }
    long n = -1;
    if (toed_oracle_convolve(img, H, W, sigma, nthreads, bufs[0], bufs[1], bufs[2], bufs[3], NULL) == 0)
        n = toed_oracle_nms(bufs[0], bufs[1], bufs[2], bufs[3], H, W, thresh, border, nthreads,
                            edges, cap, eidx, mmag, msub);
    for (int k = 0; k < 4; ++k) free(own[k]);
    return n;
}
