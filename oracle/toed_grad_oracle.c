/*
 * TEST INFRASTRUCTURE ONLY -- FP64 oracle of the differentiable TOED form.
 *
 * The reference has no differentiation.  What is restated here is the
 * reference's forward arithmetic for one edgel (cpu_toed.cpp:194-232 for the
 * nine responses at a pixel, :411-489 for the octant, interpolation and
 * parabola, :516-529 for the output row) evaluated on a FROZEN edgel set:
 * each edgel keeps its interp pixel and its octant, so the outputs
 * [x, y, theta, strength] are smooth functions of the image and of sigma
 * (with the window radii R0, R1 frozen as well).
 *
 *   toed_oracle_taps_frozen   taps at sigma for given radii, and their closed-
 *                             form d/dsigma
 *   toed_oracle_frozen_loss   L = sum_e w_e . [x_e, y_e, theta_e, s_e]
 *   toed_oracle_frozen_grad   dL/dimage and dL/dsigma by an independent route:
 *                             central differences of each edgel's output with
 *                             respect to its 17 local derivative values (the 9
 *                             responses at the edgel, Ix and Iy at the 4
 *                             interpolation neighbours), then the transposed
 *                             filters; d/dsigma of the derivative values by
 *                             central differences in sigma.
 *
 * Pinned by: central finite differences of toed_oracle_frozen_loss itself
 * (tests/test_grad_oracle.py), and the tap derivatives against finite
 * differences of the taps.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define GMAXR1 32

int toed_oracle_taps_frozen(double sigma, int R0, int R1, double *taps, double *dtaps) {
    if (!(sigma > 0.0) || R0 < 1 || R1 != R0 + 1 || R1 > GMAXR1) return -1;
    const int n = 2 * R1 + 1;
    const double s = sigma, s2 = s * s;
    const double norm = sqrt(2.0 * 3.14159265358979323846) * s;
    for (int sh = 0; sh < 2; ++sh) {
        for (int k = 0; k < n; ++k) {
            const double x = (double)(k - R1) + (sh ? 0.5 : 0.0);
            const double g = exp(-x * x / (2.0 * s2)) / norm;
            /* f_k = P_k(x, s) g,  dg/ds = g (x^2/s^3 - 1/s) */
            const double P[4] = {1.0, -x / s2, (x * x - s2) / (s2 * s2), x * (3.0 * s2 - x * x) / (s2 * s2 * s2)};
            const double dP[4] = {0.0, 2.0 * x / (s2 * s), -4.0 * x * x / (s2 * s2 * s) + 2.0 / (s2 * s),
                                  -12.0 * x / (s2 * s2 * s) + 6.0 * x * x * x / (s2 * s2 * s2 * s)};
            const double dg = g * (x * x / (s2 * s) - 1.0 / s);
            for (int f = 0; f < 4; ++f) {
                taps[(4 * sh + f) * n + k] = P[f] * g;
                if (dtaps) dtaps[(4 * sh + f) * n + k] = dP[f] * g + P[f] * dg;
            }
        }
    }
    return 0;
}

/* the nine responses at interp pixel (R, C), cpu_toed.cpp's direct 2-D sums
 * (phase (R&1, C&1); phase 0 on the R0 window) */
static void point_resp(const double *img, int H, int W, int R, int C, int R0, int R1, const double *tp, double o[9]) {
    const int n = 2 * R1 + 1;
    const int i = R >> 1, j = C >> 1, ph = ((R & 1) << 1) | (C & 1);
    const double *ux = tp + ((ph & 1) ? 4 : 0) * n;
    const double *uy = tp + ((ph & 2) ? 4 : 0) * n;
    const int rad = ph == 0 ? R0 : R1;
    double a[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int p = -rad; p <= rad; ++p) {
        const int ii = i - p;
        if (ii < 0 || ii >= H) continue;
        const int kp = p + R1;
        for (int q = -rad; q <= rad; ++q) {
            const int jj = j - q;
            if (jj < 0 || jj >= W) continue;
            const int kq = q + R1;
            const double v = img[(size_t)ii * W + jj];
            a[0] += (v * ux[n + kq]) * uy[kp];
            a[1] += (v * ux[kq]) * uy[n + kp];
            a[2] += (v * ux[2 * n + kq]) * uy[kp];
            a[3] += (v * ux[n + kq]) * uy[n + kp];
            a[4] += (v * ux[kq]) * uy[2 * n + kp];
            a[5] += (v * ux[2 * n + kq]) * uy[n + kp];
            a[6] += (v * ux[n + kq]) * uy[2 * n + kp];
            a[7] += (v * ux[3 * n + kq]) * uy[kp];
            a[8] += (v * ux[kq]) * uy[3 * n + kp];
        }
    }
    memcpy(o, a, sizeof a);
}

/* octant rules of cpu_toed.cpp:415-462 (numbering 1..8) and the neighbour
 * offsets (drow, dcol) of fp's axis/diagonal samples and fm's */
static const int OCT_NB[9][8] = {
    {0, 0, 0, 0, 0, 0, 0, 0},
    {0, 1, 1, 1, 0, -1, -1, -1},
    {1, 0, 1, 1, -1, 0, -1, -1},
    {1, 0, 1, -1, -1, 0, -1, 1},
    {0, -1, 1, -1, 0, 1, -1, 1},
    {0, -1, -1, -1, 0, 1, 1, 1},
    {-1, 0, -1, -1, 1, 0, 1, 1},
    {-1, 0, -1, 1, 1, 0, 1, -1},
    {0, 1, -1, 1, 0, -1, 1, -1},
};

static int octant_of(double gx, double gy) {
    if (gx >= 0 && gy >= 0) return gx >= gy ? 1 : 2;
    if (gx < 0 && gy >= 0) return fabs(gx) < gy ? 3 : 4;
    if (gx < 0 && gy < 0) return fabs(gx) >= fabs(gy) ? 5 : 6;
    if (gx >= 0 && gy < 0) return gx < fabs(gy) ? 7 : 8;
    return 0;
}

double toed_oracle_to_theta(double fx, double fy, double fxx, double fxy, double fyy,
                            double fxxy, double fxyy, double fxxx, double fyyy);

/* one edgel's outputs from its 17 local values: d[0..8] the nine responses at
 * the edgel, d[9 + 2k], d[10 + 2k] Ix, Iy at neighbour k (fp axis, fp diag,
 * fm axis, fm diag).  x, y are returned without the integer part C/2, R/2. */
static void edgel_out(const double *d, int oct, double out[4]) {
    const double gx = d[0], gy = d[1];
    const double f = sqrt(gx * gx + gy * gy);
    const double nx = gx / f, ny = gy / f;
    double slope;
    switch (oct) {
    case 1: case 5: slope = ny / nx; break;
    case 2: case 6: slope = nx / ny; break;
    case 3: case 7: slope = -nx / ny; break;
    default: slope = -ny / nx; break;
    }
    double M[4];
    for (int k = 0; k < 4; ++k) M[k] = sqrt(d[9 + 2 * k] * d[9 + 2 * k] + d[10 + 2 * k] * d[10 + 2 * k]);
    const double fp = M[0] * (1 - slope) + M[1] * slope;
    const double fm = M[2] * (1 - slope) + M[3] * slope;
    const double s = sqrt(1 + slope * slope);
    const double A = (fm + fp - 2 * f) / (2 * s * s);
    const double B = (fp - fm) / (2 * s);
    const double t = -B / (2 * A);
    const double maxf = A * t * t + B * t + f;
    out[0] = (t * nx - 1) / 2;
    out[1] = (t * ny - 1) / 2;
    out[2] = toed_oracle_to_theta(d[0], d[1], d[2], d[3], d[4], d[5], d[6], d[7], d[8]);
    const double sgx = maxf * nx, sgy = maxf * ny;
    out[3] = sqrt(sgx * sgx + sgy * sgy);
}

/* the 17 local values of edgel (R, C) with octant oct (0: decide from the image) */
static int edgel_values(const double *img, int H, int W, int R, int C, int R0, int R1, const double *tp,
                        int oct, double d[17], int *oct_out) {
    double c9[9];
    point_resp(img, H, W, R, C, R0, R1, tp, c9);
    memcpy(d, c9, sizeof c9);
    if (oct == 0) oct = octant_of(c9[0], c9[1]);
    if (oct == 0) return -1;
    for (int k = 0; k < 4; ++k) {
        double v[9];
        point_resp(img, H, W, R + OCT_NB[oct][2 * k], C + OCT_NB[oct][2 * k + 1], R0, R1, tp, v);
        d[9 + 2 * k] = v[0];
        d[10 + 2 * k] = v[1];
    }
    *oct_out = oct;
    return 0;
}

/* L = sum_e w_e . [x, y, theta, s]; oct[e] == 0 is decided here and stored */
double toed_oracle_frozen_loss(const double *img, int H, int W, double sigma, int R0, int R1,
                               const int64_t *eidx, long n, const double *w, int *oct, double *outs) {
    double tp[8 * (2 * GMAXR1 + 1)];
    if (toed_oracle_taps_frozen(sigma, R0, R1, tp, NULL)) return NAN;
    const int IW = 2 * W;
    double L = 0.0;
    for (long e = 0; e < n; ++e) {
        const int R = (int)(eidx[e] / IW), C = (int)(eidx[e] % IW);
        double d[17], o[4];
        int oc = oct ? oct[e] : 0;
        if (edgel_values(img, H, W, R, C, R0, R1, tp, oc, d, &oc)) continue;
        if (oct) oct[e] = oc;
        edgel_out(d, oc, o);
        o[0] += 0.5 * C;
        o[1] += 0.5 * R;
        for (int k = 0; k < 4; ++k) {
            if (outs) outs[4 * e + k] = o[k];
            if (w[4 * e + k] != 0.0) L += w[4 * e + k] * o[k];
        }
    }
    return L;
}

/* add coef * (filter of response `resp` at interp pixel (R, C)) to gimg */
static void scatter_filter(double *gimg, int H, int W, int R, int C, int resp, double coef, int R0, int R1,
                           const double *tp) {
    static const int XF[9] = {1, 0, 2, 1, 0, 2, 1, 3, 0};   /* x (q) family per response */
    static const int YF[9] = {0, 1, 0, 1, 2, 1, 2, 0, 3};   /* y (p) family */
    const int n = 2 * R1 + 1;
    const int i = R >> 1, j = C >> 1, ph = ((R & 1) << 1) | (C & 1);
    const double *X = tp + (((ph & 1) ? 4 : 0) + XF[resp]) * n;
    const double *Y = tp + (((ph & 2) ? 4 : 0) + YF[resp]) * n;
    const int rad = ph == 0 ? R0 : R1;
    for (int p = -rad; p <= rad; ++p) {
        const int ii = i - p;
        if (ii < 0 || ii >= H) continue;
        for (int q = -rad; q <= rad; ++q) {
            const int jj = j - q;
            if (jj < 0 || jj >= W) continue;
            gimg[(size_t)ii * W + jj] += coef * Y[p + R1] * X[q + R1];
        }
    }
}

int toed_oracle_frozen_grad(const double *img, int H, int W, double sigma, int R0, int R1,
                            const int64_t *eidx, long n, const double *w, const int *oct,
                            double *gimg, double *gsigma) {
    double tp[8 * (2 * GMAXR1 + 1)], tpp[8 * (2 * GMAXR1 + 1)], tpm[8 * (2 * GMAXR1 + 1)];
    const double eps_s = 1e-5;
    if (toed_oracle_taps_frozen(sigma, R0, R1, tp, NULL)) return -1;
    if (toed_oracle_taps_frozen(sigma + eps_s, R0, R1, tpp, NULL)) return -1;
    if (toed_oracle_taps_frozen(sigma - eps_s, R0, R1, tpm, NULL)) return -1;
    const int IW = 2 * W;
    memset(gimg, 0, sizeof(double) * (size_t)H * W);
    double gs = 0.0;
    for (long e = 0; e < n; ++e) {
        const int R = (int)(eidx[e] / IW), C = (int)(eidx[e] % IW);
        double d[17], dp[17], dm[17];
        int oc = oct ? oct[e] : 0;
        if (edgel_values(img, H, W, R, C, R0, R1, tp, oc, d, &oc)) continue;
        int tmp;
        edgel_values(img, H, W, R, C, R0, R1, tpp, oc, dp, &tmp);
        edgel_values(img, H, W, R, C, R0, R1, tpm, oc, dm, &tmp);
        /* adjoint of the local value k: central difference of w . out */
        for (int k = 0; k < 17; ++k) {
            const double h = 1e-6 * (fabs(d[k]) > 1e-3 ? fabs(d[k]) : 1e-3);
            double dd[17], op[4], om[4];
            memcpy(dd, d, sizeof dd);
            dd[k] = d[k] + h;
            edgel_out(dd, oc, op);
            dd[k] = d[k] - h;
            edgel_out(dd, oc, om);
            double adj = 0.0;
            for (int c = 0; c < 4; ++c) {
                double dif = op[c] - om[c];
                if (c == 2) {   /* angles: unwrap across the atan2 cut */
                    if (dif > 3.14159265358979323846) dif -= 2 * 3.14159265358979323846;
                    if (dif < -3.14159265358979323846) dif += 2 * 3.14159265358979323846;
                }
                if (isfinite(dif)) adj += w[4 * e + c] * dif / (2 * h);
            }
            if (adj == 0.0) continue;
            /* the point and response this local value belongs to */
            int PR = R, PC = C, resp = k;
            if (k >= 9) {
                const int nb = (k - 9) >> 1;
                PR = R + OCT_NB[oc][2 * nb];
                PC = C + OCT_NB[oc][2 * nb + 1];
                resp = (k - 9) & 1;
            }
            scatter_filter(gimg, H, W, PR, PC, resp, adj, R0, R1, tp);
            gs += adj * (dp[k] - dm[k]) / (2 * eps_s);
        }
    }
    if (gsigma) *gsigma = gs;
    return 0;
}
