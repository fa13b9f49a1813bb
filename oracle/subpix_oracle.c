/*
 * TEST INFRASTRUCTURE ONLY -- FP64 oracle of the subpixel third-order
 * correction path (reference: subpix_TO_correction/subpix_to_correction.cpp,
 * class SubpixelTOCorrectionCPU<double>, any interp_n).
 *
 *   estimate_orientation  :307-375  triangle smoothing (r = 4, half-sample
 *                                   symmetric boundary, :47-55), central
 *                                   differences (one-sided at the borders),
 *                                   theta = mod(atan(Oyy sgn(-Oxy) / (Oxx + 1e-5)), pi)
 *   compute_dir_from_theta :296-305 dir = (cos, sin)(theta mod pi), :216-222
 *   set_edgemap / set_orientation with interp_n = n > 0 (:226-240): bilinear
 *                         :69-90    resize (corners aligned) to the work grid
 *                                   ((h - 1) 2^n + 1) x ((w - 1) 2^n + 1); the
 *                                   steps below run on it, NMS margin 3 * 2^n
 *                                   (:627-629), x, y divided by 2^n (:694-695)
 *   nms_token(margin 3)   :413-519  octant NMS on the map along dir, parabola,
 *                                   |A| < 1e-16 rejected, strength |max_f|
 *   run                   :623-736  Hx = E dirx, Hy = E diry; 8 circular
 *                                   separable convolutions (:377-411); Keys
 *                                   cubic (a = -1/2) sampling with clamped
 *                                   borders (:57-67, :521-556) of 10 planes at
 *                                   each edgel; F = (Fx, Fy); orientation
 *                                   atan2(Fx, -Fy)/|F| with the fallback
 *                                   atan2(dirx, -diry) when |F| < 1e-12;
 *                                   isolated-edgel removal (:558-621); map
 *                                   scatter, last writer wins (:708-729).
 *
 * Pinned by golden vectors of the reference's own subpix_to_correction.cpp
 * compiled from its sources (oracle/_ref, tests/golden) and by the checks in
 * tests/test_subpix_oracle.py (circular convolution = FFT convolution, Keys
 * cubic reproduces cubics, orientation of straight ridges).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

static const double SP_PI = 3.14159265358979323846;

static int wrap_idx(int i, int n) {
    int r = i % n;
    if (r < 0) r += n;
    return r;
}

static int wrap_sym(int i, int n) {
    if (n <= 1) return 0;
    while (i < 0 || i >= n) {
        if (i < 0) i = -i - 1;
        else i = 2 * n - i - 1;
    }
    return i;
}

static double wrap_mod_pi(double a) {
    double r = fmod(a, SP_PI);
    if (r < 0) r += SP_PI;
    return r;
}

static double cubic_w(double s) {
    const double a = -0.5;
    const double x = fabs(s);
    if (x <= 1.0) return ((a + 2) * x - (a + 3)) * x * x + 1;
    if (x < 2.0) return ((a * x - 5 * a) * x + 8 * a) * x - 4 * a;
    return 0.0;
}

/* S1 taps: G, G', G'' of radius ceil(4 sigma) at integer offsets (:194-214) */
int toed_subpix_taps(double sigma, int *rad, double *G, double *dG, double *d2G) {
    if (!(sigma > 0)) return -1;
    const int r = (int)ceil(sigma * 4);
    const double s2 = sigma * sigma, s3 = s2 * sigma, s5 = s2 * s3, nrm = sqrt(2.0 * SP_PI);
    for (int k = -r; k <= r; ++k) {
        const double e = exp(-((double)k * k) / (2 * s2));
        G[k + r] = e / (nrm * sigma);
        dG[k + r] = -k * e / (nrm * s3);
        d2G[k + r] = ((double)k * k - s2) * e / (nrm * s5);
    }
    *rad = r;
    return 0;
}

/* S2 */
void toed_subpix_estimate_orientation(const double *E, int h, int w, double *theta) {
    const int r = 4;
    double tri[9], tsum = 0;
    for (int k = -r; k <= r; ++k) { tri[k + r] = r + 1 - abs(k); tsum += tri[k + r]; }
    for (int k = 0; k < 9; ++k) tri[k] /= tsum;
    double *tmp = (double *)malloc(sizeof(double) * (size_t)h * w);
    double *sm = (double *)malloc(sizeof(double) * (size_t)h * w);
    double *Ox = (double *)malloc(sizeof(double) * (size_t)h * w);
    double *Oy = (double *)malloc(sizeof(double) * (size_t)h * w);
    for (int i = 0; i < h; ++i)
        for (int j = 0; j < w; ++j) {
            double acc = 0;
            for (int q = -r; q <= r; ++q) acc += E[(size_t)i * w + wrap_sym(j - q, w)] * tri[q + r];
            tmp[(size_t)i * w + j] = acc;
        }
    for (int i = 0; i < h; ++i)
        for (int j = 0; j < w; ++j) {
            double acc = 0;
            for (int p = -r; p <= r; ++p) acc += tmp[(size_t)wrap_sym(i - p, h) * w + j] * tri[p + r];
            sm[(size_t)i * w + j] = acc;
        }
#define DX(a, i, j) ((j) == 0 ? (a)[(size_t)(i) * w + 1] - (a)[(size_t)(i) * w] : \
                     (j) == w - 1 ? (a)[(size_t)(i) * w + w - 1] - (a)[(size_t)(i) * w + w - 2] : \
                     0.5 * ((a)[(size_t)(i) * w + (j) + 1] - (a)[(size_t)(i) * w + (j) - 1]))
#define DY(a, i, j) ((i) == 0 ? (a)[(size_t)w + (j)] - (a)[(j)] : \
                     (i) == h - 1 ? (a)[(size_t)(h - 1) * w + (j)] - (a)[(size_t)(h - 2) * w + (j)] : \
                     0.5 * ((a)[(size_t)((i) + 1) * w + (j)] - (a)[(size_t)((i) - 1) * w + (j)]))
    for (int i = 0; i < h; ++i)
        for (int j = 0; j < w; ++j) {
            Ox[(size_t)i * w + j] = DX(sm, i, j);
            Oy[(size_t)i * w + j] = DY(sm, i, j);
        }
    for (int i = 0; i < h; ++i)
        for (int j = 0; j < w; ++j) {
            const double oxx = DX(Ox, i, j), oxy = DX(Oy, i, j), oyy = DY(Oy, i, j);
            const double sgn = oxy < 0 ? 1.0 : (oxy > 0 ? -1.0 : 0.0);
            theta[(size_t)i * w + j] = wrap_mod_pi(atan(oyy * sgn / (oxx + 1e-5)));
        }
#undef DX
#undef DY
    free(tmp); free(sm); free(Ox); free(Oy);
}

/* S5: circular separable convolution, x pass then y pass (:377-411) */
static void conv_circ(const double *src, double *dst, double *tmp, int h, int w, int rad, const double *kx,
                      const double *ky) {
#pragma omp parallel for schedule(static)
    for (int i = 0; i < h; ++i)
        for (int j = 0; j < w; ++j) {
            double acc = 0;
            for (int q = -rad; q <= rad; ++q) acc += src[(size_t)i * w + wrap_idx(j - q, w)] * kx[q + rad];
            tmp[(size_t)i * w + j] = acc;
        }
#pragma omp parallel for schedule(static)
    for (int i = 0; i < h; ++i)
        for (int j = 0; j < w; ++j) {
            double acc = 0;
            for (int p = -rad; p <= rad; ++p) acc += tmp[(size_t)wrap_idx(i - p, h) * w + j] * ky[p + rad];
            dst[(size_t)i * w + j] = acc;
        }
}

void toed_subpix_conv_circ(const double *src, double *dst, int h, int w, int rad, const double *kx, const double *ky) {
    double *tmp = (double *)malloc(sizeof(double) * (size_t)h * w);
    conv_circ(src, dst, tmp, h, w, rad, kx, ky);
    free(tmp);
}

/* S6: Keys cubic sample with clamped borders (:521-556) */
double toed_subpix_interp_cubic(const double *src, int h, int w, double x, double y) {
    const int x0 = (int)floor(x), y0 = (int)floor(y);
    const double fx = x - x0, fy = y - y0;
    const double wx[4] = {cubic_w(fx + 1), cubic_w(fx), cubic_w(fx - 1), cubic_w(fx - 2)};
    const double wy[4] = {cubic_w(fy + 1), cubic_w(fy), cubic_w(fy - 1), cubic_w(fy - 2)};
    double acc = 0;
    for (int p = 0; p < 4; ++p) {
        int ii = y0 - 1 + p;
        ii = ii < 0 ? 0 : (ii >= h ? h - 1 : ii);
        double row = 0;
        for (int q = 0; q < 4; ++q) {
            int jj = x0 - 1 + q;
            jj = jj < 0 ? 0 : (jj >= w ? w - 1 : jj);
            row += src[(size_t)ii * w + jj] * wx[q];
        }
        acc += row * wy[p];
    }
    return acc;
}

/*
 * Bilinear resize with the corners aligned (:69-90): work pixel (i, j) samples
 * the map at (i (sh - 1) / (dh - 1), j (sw - 1) / (dw - 1)), the right and
 * lower neighbours clamped to the last row / column, in the reference's
 * operation order (built with -ffp-contract=off).  f32: each result rounded to
 * float, as the GPU stores its work maps.
 */
void toed_subpix_resize(const double *src, int sh, int sw, double *dst, int dh, int dw, int f32) {
    for (int i = 0; i < dh; ++i) {
        const double y = dh == 1 ? 0.0 : (double)i * (sh - 1) / (double)(dh - 1);
        const int ya = (int)floor(y), yb = ya + 1 < sh ? ya + 1 : sh - 1;
        const double fy = y - ya;
        const double *ra = src + (size_t)ya * sw, *rb = src + (size_t)yb * sw;
        for (int j = 0; j < dw; ++j) {
            const double x = dw == 1 ? 0.0 : (double)j * (sw - 1) / (double)(dw - 1);
            const int xa = (int)floor(x), xb = xa + 1 < sw ? xa + 1 : sw - 1;
            const double fx = x - xa;
            const double v = (1 - fy) * ((1 - fx) * ra[xa] + fx * ra[xb]) + fy * ((1 - fx) * rb[xa] + fx * rb[xb]);
            dst[(size_t)i * dw + j] = f32 ? (double)(float)v : v;
        }
    }
}

static long subpix_work(const double *E, int h, int w, const double *theta_in, double thresh, double sigma,
                        double nbr_radius, double *out, long cap, int64_t *nms_idx, double *TO_E, double *TO_O,
                        int margin, double scale, int ih, int iw);

/*
 * Whole path.  E: ih x iw map; theta: given orientation or NULL (estimated).
 * interp_n = n > 0: E and theta are resized to the work grid
 * ((ih - 1) 2^n + 1) x ((iw - 1) 2^n + 1) first (set_edgemap / set_orientation,
 * :226-240; work_f32 rounds the resized maps to float); the NMS margin is
 * 3 * 2^n (:627-629) and x, y are divided by 2^n (:694-695).
 * out: N x 4 [x, y, orientation, strength] after isolated-edgel removal;
 * nms_idx (optional): the work-grid pixel i*w'+j of each output edgel; TO_E /
 * TO_O (optional) ih x iw maps.  Returns N (rows beyond cap counted, not written).
 */
long toed_subpix_oracle_n(const double *E, int ih, int iw, const double *theta_in, double thresh, double sigma,
                          double nbr_radius, int nthreads, int interp_n, int work_f32, double *out, long cap,
                          int64_t *nms_idx, double *TO_E, double *TO_O) {
    if (ih <= 0 || iw <= 0 || !E || !(sigma > 0) || interp_n < 0 || interp_n > 8) return -1;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
    if (interp_n == 0 && !work_f32)
        return subpix_work(E, ih, iw, theta_in, thresh, sigma, nbr_radius, out, cap, nms_idx, TO_E, TO_O, 3, 1.0, ih,
                           iw);
    const int s = 1 << interp_n;
    const int h = (ih - 1) * s + 1, w = (iw - 1) * s + 1;
    double *Ew = (double *)malloc(sizeof(double) * (size_t)h * w);
    double *Tw = theta_in ? (double *)malloc(sizeof(double) * (size_t)h * w) : NULL;
    toed_subpix_resize(E, ih, iw, Ew, h, w, work_f32);
    if (Tw) toed_subpix_resize(theta_in, ih, iw, Tw, h, w, work_f32);
    const long N = subpix_work(Ew, h, w, Tw, thresh, sigma, nbr_radius, out, cap, nms_idx, TO_E, TO_O, 3 * s,
                               (double)s, ih, iw);
    free(Ew);
    free(Tw);
    return N;
}

long toed_subpix_oracle(const double *E, int h, int w, const double *theta_in, double thresh, double sigma,
                        double nbr_radius, int nthreads, double *out, long cap, int64_t *nms_idx, double *TO_E,
                        double *TO_O) {
    return toed_subpix_oracle_n(E, h, w, theta_in, thresh, sigma, nbr_radius, nthreads, 0, 0, out, cap, nms_idx,
                                TO_E, TO_O);
}

/* The path on an h x w work grid; TO maps ih x iw (image grid), positions
 * divided by `scale` (a power of two: exact) before isolation and output. */
static long subpix_work(const double *E, int h, int w, const double *theta_in, double thresh, double sigma,
                        double nbr_radius, double *out, long cap, int64_t *nms_idx, double *TO_E, double *TO_O,
                        int margin, double scale, int ih, int iw) {
    const size_t n = (size_t)h * w;
    int rad;
    double G[129], dG[129], d2G[129];
    if ((int)ceil(sigma * 4) > 64 || toed_subpix_taps(sigma, &rad, G, dG, d2G)) return -1;
    double *theta = (double *)malloc(sizeof(double) * n);
    double *dirx = (double *)malloc(sizeof(double) * n), *diry = (double *)malloc(sizeof(double) * n);
    double *P[10];
    for (int k = 0; k < 10; ++k) P[k] = (double *)malloc(sizeof(double) * n);
    double *tmp = (double *)malloc(sizeof(double) * n);
    double *nx = (double *)calloc(n, sizeof(double)), *ny = (double *)calloc(n, sizeof(double));
    double *nm = (double *)calloc(n, sizeof(double));
    unsigned char *nv = (unsigned char *)calloc(n, 1);
    if (theta_in) memcpy(theta, theta_in, sizeof(double) * n);
    else toed_subpix_estimate_orientation(E, h, w, theta);
    for (size_t k = 0; k < n; ++k) {
        const double a = wrap_mod_pi(theta[k]);
        dirx[k] = cos(a);
        diry[k] = sin(a);
    }
    /* S4: nms_token(margin): rows/cols [margin + 1, n - margin - 2) */
    const int j0 = margin + 1, j1 = w - (margin + 2), i0 = margin + 1, i1 = h - (margin + 2);
#define EM(a, b) E[(size_t)(a) * w + (b)]
#pragma omp parallel for schedule(dynamic)
    for (int j = j0; j < j1; ++j) {
        for (int i = i0; i < i1; ++i) {
            if (EM(i, j) <= thresh) continue;
            const double gx = dirx[(size_t)i * w + j], gy = diry[(size_t)i * w + j];
            if (fabs(gx) < 1e-5 && fabs(gy) < 1e-5) continue;
            const double md = sqrt(gx * gx + gy * gy);
            const double ndx = gx / md, ndy = gy / md, f = EM(i, j);
            double slope, fp, fm;
            if (gx >= 0 && gy >= 0) {
                if (gx >= gy) { slope = ndy / ndx; fp = EM(i, j + 1) * (1 - slope) + EM(i + 1, j + 1) * slope; fm = EM(i, j - 1) * (1 - slope) + EM(i - 1, j - 1) * slope; }
                else { slope = ndx / ndy; fp = EM(i + 1, j) * (1 - slope) + EM(i + 1, j + 1) * slope; fm = EM(i - 1, j) * (1 - slope) + EM(i - 1, j - 1) * slope; }
            } else if (gx < 0 && gy >= 0) {
                if (fabs(gx) < gy) { slope = -ndx / ndy; fp = EM(i + 1, j) * (1 - slope) + EM(i + 1, j - 1) * slope; fm = EM(i - 1, j) * (1 - slope) + EM(i - 1, j + 1) * slope; }
                else { slope = -ndy / ndx; fp = EM(i, j - 1) * (1 - slope) + EM(i + 1, j - 1) * slope; fm = EM(i, j + 1) * (1 - slope) + EM(i - 1, j + 1) * slope; }
            } else if (gx < 0 && gy < 0) {
                if (fabs(gx) >= fabs(gy)) { slope = ndy / ndx; fp = EM(i, j - 1) * (1 - slope) + EM(i - 1, j - 1) * slope; fm = EM(i, j + 1) * (1 - slope) + EM(i + 1, j + 1) * slope; }
                else { slope = ndx / ndy; fp = EM(i - 1, j) * (1 - slope) + EM(i - 1, j - 1) * slope; fm = EM(i + 1, j) * (1 - slope) + EM(i + 1, j + 1) * slope; }
            } else {
                if (gx < fabs(gy)) { slope = -ndx / ndy; fp = EM(i - 1, j) * (1 - slope) + EM(i - 1, j + 1) * slope; fm = EM(i + 1, j) * (1 - slope) + EM(i + 1, j - 1) * slope; }
                else { slope = -ndy / ndx; fp = EM(i, j + 1) * (1 - slope) + EM(i - 1, j + 1) * slope; fm = EM(i, j - 1) * (1 - slope) + EM(i + 1, j - 1) * slope; }
            }
            if (!((f > fm && f > fp) || (f > fm && f >= fp) || (f >= fm && f > fp))) continue;
            const double s = sqrt(1 + slope * slope);
            const double A = (fm + fp - 2 * f) / (2 * s * s);
            const double B = (fp - fm) / (2 * s);
            if (fabs(A) < 1e-16) continue;
            const double ss = -B / (2 * A);
            const double maxf = A * ss * ss + B * ss + f;
            if (fabs(ss) <= sqrt(2.0)) {
                nx[(size_t)i * w + j] = j + ss * ndx;
                ny[(size_t)i * w + j] = i + ss * ndy;
                nm[(size_t)i * w + j] = fabs(maxf);
                nv[(size_t)i * w + j] = 1;
            }
        }
    }
#undef EM
    /* S5: P[0..9] = Px, Py, Pxx, Pyy, Pxy, Hx, Hxx, Hxy, Hyy, Hy */
    double *Hxm = P[5], *Hym = P[9];
    for (size_t k = 0; k < n; ++k) { Hxm[k] = E[k] * dirx[k]; Hym[k] = E[k] * diry[k]; }
    conv_circ(E, P[0], tmp, h, w, rad, dG, G);
    conv_circ(E, P[1], tmp, h, w, rad, G, dG);
    conv_circ(E, P[2], tmp, h, w, rad, d2G, G);
    conv_circ(E, P[4], tmp, h, w, rad, dG, dG);
    conv_circ(E, P[3], tmp, h, w, rad, G, d2G);
    conv_circ(Hxm, P[6], tmp, h, w, rad, dG, G);
    conv_circ(Hym, P[8], tmp, h, w, rad, G, dG);
    conv_circ(Hym, P[7], tmp, h, w, rad, dG, G);
    /* S6: per edgel in raster order */
    long ne = 0;
    for (size_t k = 0; k < n; ++k) ne += nv[k];
    double *ex = (double *)malloc(sizeof(double) * (ne + 1) * 4);
    int64_t *eidx = (int64_t *)malloc(sizeof(int64_t) * (ne + 1));
    long m = 0;
    for (int i = 0; i < h; ++i)
        for (int j = 0; j < w; ++j) {
            const size_t o = (size_t)i * w + j;
            if (!nv[o]) continue;
            const double sx = nx[o], sy = ny[o];
            double v[10];
            for (int k = 0; k < 10; ++k) v[k] = toed_subpix_interp_cubic(P[k], h, w, sx, sy);
            const double Px_ = v[0], Py_ = v[1], Pxx_ = v[2], Pyy_ = v[3], Pxy_ = v[4];
            const double Hx_ = v[5], Hxx_ = v[6], Hxy_ = v[7], Hyy_ = v[8], Hy_ = v[9];
            double Fx = Px_ * Hxx_ + Pxx_ * Hx_ + Py_ * Hxy_ + Pxy_ * Hy_;
            double Fy = Px_ * Hxy_ + Pxy_ * Hx_ + Py_ * Hyy_ + Pyy_ * Hy_;
            const double Fm = sqrt(Fx * Fx + Fy * Fy);
            double ori;
            if (Fm < 1e-12) ori = atan2(dirx[o], -diry[o]);
            else { Fx /= Fm; Fy /= Fm; ori = atan2(Fx, -Fy); }
            ex[4 * m + 0] = sx / scale; ex[4 * m + 1] = sy / scale; ex[4 * m + 2] = ori; ex[4 * m + 3] = nm[o];
            eidx[m] = (int64_t)o;
            ++m;
        }
    /* S7: keep edgels with another edgel within nbr_radius (grid of cells of size r) */
    char *keep = (char *)malloc(ne + 1);
    if (nbr_radius > 0 && ne >= 2) {
        const double r2 = nbr_radius * nbr_radius, inv = 1.0 / nbr_radius;
        /* brute force over cells: sort by cell via a simple spatial hash on a dense cell grid */
        const int cw = (int)floor((w + 2) * inv) + 3, ch = (int)floor((h + 2) * inv) + 3;
        int *head = (int *)malloc(sizeof(int) * (size_t)cw * ch), *next = (int *)malloc(sizeof(int) * ne);
        int *ccx = (int *)malloc(sizeof(int) * ne), *ccy = (int *)malloc(sizeof(int) * ne);
        for (size_t k = 0; k < (size_t)cw * ch; ++k) head[k] = -1;
        for (long a = 0; a < ne; ++a) {
            ccx[a] = (int)floor(ex[4 * a] * inv);
            ccy[a] = (int)floor(ex[4 * a + 1] * inv);
            const int cx = ccx[a] + 1, cy = ccy[a] + 1;
            next[a] = head[(size_t)cy * cw + cx];
            head[(size_t)cy * cw + cx] = (int)a;
        }
        for (long a = 0; a < ne; ++a) {
            int found = 0;
            for (int dy = -1; dy <= 1 && !found; ++dy)
                for (int dx = -1; dx <= 1 && !found; ++dx) {
                    const int cx = ccx[a] + 1 + dx, cy = ccy[a] + 1 + dy;
                    if (cx < 0 || cy < 0 || cx >= cw || cy >= ch) continue;
                    for (int b = head[(size_t)cy * cw + cx]; b >= 0; b = next[b]) {
                        if (b == a) continue;
                        const double ddx = ex[4 * a] - ex[4 * b], ddy = ex[4 * a + 1] - ex[4 * b + 1];
                        if (ddx * ddx + ddy * ddy <= r2) { found = 1; break; }
                    }
                }
            keep[a] = (char)found;
        }
        free(head); free(next); free(ccx); free(ccy);
    } else {
        for (long a = 0; a < ne; ++a) keep[a] = 1;
    }
    /* S8 */
    if (TO_E) memset(TO_E, 0, sizeof(double) * (size_t)ih * iw);
    if (TO_O) memset(TO_O, 0, sizeof(double) * (size_t)ih * iw);
    long N = 0;
    for (long a = 0; a < ne; ++a) {
        if (!keep[a]) continue;
        if (TO_E || TO_O) {
            int X = (int)llround(ex[4 * a]), Y = (int)llround(ex[4 * a + 1]);
            X = X < 0 ? 0 : (X > iw - 1 ? iw - 1 : X);
            Y = Y < 0 ? 0 : (Y > ih - 1 ? ih - 1 : Y);
            if (TO_E) TO_E[(size_t)Y * iw + X] = ex[4 * a + 3];
            if (TO_O) TO_O[(size_t)Y * iw + X] = ex[4 * a + 2];
        }
        if (N < cap && out) memcpy(out + 4 * N, ex + 4 * a, 4 * sizeof(double));
        if (N < cap && nms_idx) nms_idx[N] = eidx[a];
        ++N;
    }
    free(keep); free(ex); free(eidx);
    free(theta); free(dirx); free(diry); free(tmp); free(nx); free(ny); free(nm); free(nv);
    for (int k = 0; k < 10; ++k) free(P[k]);
    return N;
}

/* ---------------------------------------------------------------------------
 * Differentiable form (no reference counterpart): the path as a function of
 * the map E with every discrete decision frozen at a forward point
 *   - the edgel set (map pixels idx[0..n)), the NMS octant of each edgel,
 *   - per pixel: the sign of t = Oyy sgn / (Oxx + 1e-5) (which half-turn the
 *     direction takes) and sgn = sign(-Oxy),
 *   - per edgel: the cubic stencil base (floor of the subpixel position) and
 *     the |F| < 1e-12 fallback.
 * toed_subpix_frozen_state() extracts the state from E; toed_subpix_frozen_loss()
 * evaluates L = sum_e w_e . [x, y, orientation, strength] with it (FP64).
 * With a given orientation map theta (set_orientation, :226-232) the direction
 * (cos, sin)(theta mod pi) is a constant input and tneg / sxy are unused.
 * Central differences of the frozen loss pin the GPU adjoint.
 */
static void sp_smooth(const double *E, int h, int w, double *S) {
    const int r = 4;
    double tri[9], tsum = 0;
    for (int k = -r; k <= r; ++k) { tri[k + r] = r + 1 - abs(k); tsum += tri[k + r]; }
    for (int k = 0; k < 9; ++k) tri[k] /= tsum;
    double *tmp = (double *)malloc(sizeof(double) * (size_t)h * w);
    for (int i = 0; i < h; ++i)
        for (int j = 0; j < w; ++j) {
            double acc = 0;
            for (int q = -r; q <= r; ++q) acc += E[(size_t)i * w + wrap_sym(j - q, w)] * tri[q + r];
            tmp[(size_t)i * w + j] = acc;
        }
    for (int i = 0; i < h; ++i)
        for (int j = 0; j < w; ++j) {
            double acc = 0;
            for (int p = -r; p <= r; ++p) acc += tmp[(size_t)wrap_sym(i - p, h) * w + j] * tri[p + r];
            S[(size_t)i * w + j] = acc;
        }
    free(tmp);
}

static double sp_dx(const double *a, int w, int i, int j) {
    if (j == 0) return a[(size_t)i * w + 1] - a[(size_t)i * w];
    if (j == w - 1) return a[(size_t)i * w + w - 1] - a[(size_t)i * w + w - 2];
    return 0.5 * (a[(size_t)i * w + j + 1] - a[(size_t)i * w + j - 1]);
}
static double sp_dy(const double *a, int h, int w, int i, int j) {
    if (i == 0) return a[(size_t)w + j] - a[j];
    if (i == h - 1) return a[(size_t)(h - 1) * w + j] - a[(size_t)(h - 2) * w + j];
    return 0.5 * (a[(size_t)(i + 1) * w + j] - a[(size_t)(i - 1) * w + j]);
}

/* Oxx + 1e-5, Oyy, Oxy at every pixel (estimate_orientation :346-367) */
static void sp_second(const double *E, int h, int w, double *Axx, double *Ayy, double *Axy) {
    const size_t n = (size_t)h * w;
    double *S = (double *)malloc(sizeof(double) * n), *Ox = (double *)malloc(sizeof(double) * n),
           *Oy = (double *)malloc(sizeof(double) * n);
    sp_smooth(E, h, w, S);
    for (int i = 0; i < h; ++i)
        for (int j = 0; j < w; ++j) {
            Ox[(size_t)i * w + j] = sp_dx(S, w, i, j);
            Oy[(size_t)i * w + j] = sp_dy(S, h, w, i, j);
        }
    for (int i = 0; i < h; ++i)
        for (int j = 0; j < w; ++j) {
            Axx[(size_t)i * w + j] = sp_dx(Ox, w, i, j) + 1e-5;
            if (Axy) Axy[(size_t)i * w + j] = sp_dx(Oy, w, i, j);
            Ayy[(size_t)i * w + j] = sp_dy(Oy, h, w, i, j);
        }
    free(S); free(Ox); free(Oy);
}

/* octant index 0..7 of a direction, the branch order of nms_token (:452-498) */
static int sp_octant(double gx, double gy) {
    if (gx >= 0 && gy >= 0) return gx >= gy ? 0 : 1;
    if (gx < 0 && gy >= 0) return fabs(gx) < gy ? 2 : 3;
    if (gx < 0 && gy < 0) return fabs(gx) >= fabs(gy) ? 4 : 5;
    return gx < fabs(gy) ? 6 : 7;
}

/* slope and the two interpolated neighbours of octant `oc` (:452-498) */
static void sp_nms_interp(const double *E, int w, int i, int j, double nx, double ny, int oc, double *slope,
                          double *fp, double *fm) {
#define EE(a, b) E[(size_t)(a) * w + (b)]
    double s;
    switch (oc) {
    case 0: s = ny / nx; *fp = EE(i, j + 1) * (1 - s) + EE(i + 1, j + 1) * s; *fm = EE(i, j - 1) * (1 - s) + EE(i - 1, j - 1) * s; break;
    case 1: s = nx / ny; *fp = EE(i + 1, j) * (1 - s) + EE(i + 1, j + 1) * s; *fm = EE(i - 1, j) * (1 - s) + EE(i - 1, j - 1) * s; break;
    case 2: s = -nx / ny; *fp = EE(i + 1, j) * (1 - s) + EE(i + 1, j - 1) * s; *fm = EE(i - 1, j) * (1 - s) + EE(i - 1, j + 1) * s; break;
    case 3: s = -ny / nx; *fp = EE(i, j - 1) * (1 - s) + EE(i + 1, j - 1) * s; *fm = EE(i, j + 1) * (1 - s) + EE(i - 1, j + 1) * s; break;
    case 4: s = ny / nx; *fp = EE(i, j - 1) * (1 - s) + EE(i - 1, j - 1) * s; *fm = EE(i, j + 1) * (1 - s) + EE(i + 1, j + 1) * s; break;
    case 5: s = nx / ny; *fp = EE(i - 1, j) * (1 - s) + EE(i - 1, j - 1) * s; *fm = EE(i + 1, j) * (1 - s) + EE(i + 1, j + 1) * s; break;
    case 6: s = -nx / ny; *fp = EE(i - 1, j) * (1 - s) + EE(i - 1, j + 1) * s; *fm = EE(i + 1, j) * (1 - s) + EE(i + 1, j - 1) * s; break;
    default: s = -ny / nx; *fp = EE(i, j + 1) * (1 - s) + EE(i - 1, j + 1) * s; *fm = EE(i, j - 1) * (1 - s) + EE(i + 1, j - 1) * s; break;
    }
    *slope = s;
#undef EE
}

/* frozen state at E: tneg[p] = 1 if t < 0 at pixel p, sxy[p] = sgn(-Oxy) (-1, 0, 1),
 * per edgel: octant, stencil base (floor x, floor y), fallback flag. */
int toed_subpix_frozen_state(const double *E, int h, int w, const double *theta_in, const int64_t *idx, long n,
                             signed char *tneg, signed char *sxy, int *oct, int *base, int *fallback, double sigma) {
    const size_t np = (size_t)h * w;
    (void)sigma;
    if (!theta_in) {
        double *Axx = (double *)malloc(sizeof(double) * np), *Ayy = (double *)malloc(sizeof(double) * np),
               *Axy = (double *)malloc(sizeof(double) * np);
        sp_second(E, h, w, Axx, Ayy, Axy);
        for (size_t p = 0; p < np; ++p) {
            const double sg = Axy[p] < 0 ? 1.0 : (Axy[p] > 0 ? -1.0 : 0.0);
            sxy[p] = (signed char)sg;
            tneg[p] = (Ayy[p] * sg / Axx[p]) < 0.0 ? 1 : 0;
        }
        free(Axx); free(Ayy); free(Axy);
    } else {
        memset(sxy, 0, np);
        memset(tneg, 0, np);
    }
    /* octants and cubic stencil bases at the forward point (:446-516, :521-530) */
    double *th = (double *)malloc(sizeof(double) * np);
    if (theta_in) memcpy(th, theta_in, sizeof(double) * np);
    else toed_subpix_estimate_orientation(E, h, w, th);
    for (long e = 0; e < n; ++e) {
        const size_t p = (size_t)idx[e];
        const int i = (int)(p / w), j = (int)(p % w);
        const double a = wrap_mod_pi(th[p]), gx = cos(a), gy = sin(a);
        const double md = sqrt(gx * gx + gy * gy), nx = gx / md, ny = gy / md;
        double slope, fp, fm;
        oct[e] = sp_octant(gx, gy);
        sp_nms_interp(E, w, i, j, nx, ny, oct[e], &slope, &fp, &fm);
        const double f = E[p], s = sqrt(1 + slope * slope);
        const double A = (fm + fp - 2 * f) / (2 * s * s), B = (fp - fm) / (2 * s), ss = -B / (2 * A);
        base[2 * e] = (int)floor(j + ss * nx);
        base[2 * e + 1] = (int)floor(i + ss * ny);
        fallback[e] = 0;   /* |F| < 1e-12: not met on the maps this is used with */
    }
    free(th);
    return 0;
}

/* The frozen loss.  w4: n x 4 weights (or one row of 4 when per_edgel == 0). */
double toed_subpix_frozen_loss(const double *E, int h, int w, const double *theta_in, double sigma,
                               const int64_t *idx, long n, const double *w4, int per_edgel, const signed char *tneg,
                               const signed char *sxy, const int *oct, const int *base, const int *fallback,
                               double *outs) {
    const size_t np = (size_t)h * w;
    int rad;
    double G[129], dG[129], d2G[129];
    if (toed_subpix_taps(sigma, &rad, G, dG, d2G)) return NAN;
    double *dxp = (double *)malloc(sizeof(double) * np), *dyp = (double *)malloc(sizeof(double) * np);
    if (theta_in) {
        /* given orientation: a constant direction (:296-305) */
        for (size_t p = 0; p < np; ++p) {
            const double a = wrap_mod_pi(theta_in[p]);
            dxp[p] = cos(a);
            dyp[p] = sin(a);
        }
    } else {
        double *Axx = (double *)malloc(sizeof(double) * np), *Ayy = (double *)malloc(sizeof(double) * np);
        sp_second(E, h, w, Axx, Ayy, NULL);
        for (size_t p = 0; p < np; ++p) {
            const double t = Ayy[p] * sxy[p] / Axx[p];
            const double inv = 1.0 / sqrt(1.0 + t * t), sg = tneg[p] ? -1.0 : 1.0;
            dxp[p] = sg * inv;
            dyp[p] = sg * t * inv;
        }
        free(Axx); free(Ayy);
    }
    double *P[10], *tmp = (double *)malloc(sizeof(double) * np);
    for (int k = 0; k < 10; ++k) P[k] = (double *)malloc(sizeof(double) * np);
    for (size_t p = 0; p < np; ++p) { P[5][p] = E[p] * dxp[p]; P[9][p] = E[p] * dyp[p]; }
    conv_circ(E, P[0], tmp, h, w, rad, dG, G);
    conv_circ(E, P[1], tmp, h, w, rad, G, dG);
    conv_circ(E, P[2], tmp, h, w, rad, d2G, G);
    conv_circ(E, P[4], tmp, h, w, rad, dG, dG);
    conv_circ(E, P[3], tmp, h, w, rad, G, d2G);
    conv_circ(P[5], P[6], tmp, h, w, rad, dG, G);
    conv_circ(P[9], P[8], tmp, h, w, rad, G, dG);
    conv_circ(P[9], P[7], tmp, h, w, rad, dG, G);
    double L = 0.0;
    for (long e = 0; e < n; ++e) {
        const int i = (int)(idx[e] / w), j = (int)(idx[e] % w);
        const double gx = dxp[(size_t)idx[e]], gy = dyp[(size_t)idx[e]];
        const double md = sqrt(gx * gx + gy * gy), nx = gx / md, ny = gy / md;
        double slope, fp, fm;
        sp_nms_interp(E, w, i, j, nx, ny, oct[e], &slope, &fp, &fm);
        const double f = E[(size_t)idx[e]];
        const double s = sqrt(1 + slope * slope);
        const double A = (fm + fp - 2 * f) / (2 * s * s), B = (fp - fm) / (2 * s);
        const double ss = -B / (2 * A), mf = A * ss * ss + B * ss + f;
        const double x = j + ss * nx, y = i + ss * ny;
        const int x0 = base[2 * e], y0 = base[2 * e + 1];
        const double fx = x - x0, fy = y - y0;
        const double wx[4] = {cubic_w(fx + 1), cubic_w(fx), cubic_w(fx - 1), cubic_w(fx - 2)};
        const double wy[4] = {cubic_w(fy + 1), cubic_w(fy), cubic_w(fy - 1), cubic_w(fy - 2)};
        double v[10];
        for (int k = 0; k < 10; ++k) {
            double acc = 0;
            for (int a = 0; a < 4; ++a) {
                int ii = y0 - 1 + a;
                ii = ii < 0 ? 0 : (ii >= h ? h - 1 : ii);
                double row = 0;
                for (int b = 0; b < 4; ++b) {
                    int jj = x0 - 1 + b;
                    jj = jj < 0 ? 0 : (jj >= w ? w - 1 : jj);
                    row += P[k][(size_t)ii * w + jj] * wx[b];
                }
                acc += row * wy[a];
            }
            v[k] = acc;
        }
        const double Fx = v[0] * v[6] + v[2] * v[5] + v[1] * v[7] + v[4] * v[9];
        const double Fy = v[0] * v[7] + v[4] * v[5] + v[1] * v[8] + v[3] * v[9];
        const double ori = fallback[e] ? atan2(gx, -gy) : atan2(Fx, -Fy);
        const double o4[4] = {x, y, ori, fabs(mf)};
        const double *we = w4 + (per_edgel ? 4 * e : 0);
        for (int k = 0; k < 4; ++k) {
            L += we[k] * o4[k];
            if (outs) outs[4 * e + k] = o4[k];
        }
    }
    for (int k = 0; k < 10; ++k) free(P[k]);
    free(tmp); free(dxp); free(dyp);
    return L;
}
