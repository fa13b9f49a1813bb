/*
 * Deterministic synthetic inputs for TOED tests and benchmarks.
 *
 * Two generators, both driven by a SplitMix64 integer RNG so that any host
 * produces bit-identical images from a seed:
 *
 *   toed_scene_u8      grey-level scene: uniform background ~U[40,200], then
 *                      24*k filled shapes painted in order (even index: disk,
 *                      odd index: convex polygon of 3-6 vertices on an ellipse
 *                      inscribed in a random box), rendered with 4x4
 *                      supersampling and box-averaged, plus N(0, 2^2) noise,
 *                      rounded and clamped to [0,255].  k = max(1, H / 540)
 *                      scales both the shape count and the shape sizes
 *                      (480x752 -> 24 shapes, 2160x3840 -> 96 shapes, x4 size).
 *
 *   toed_bezier_map_f32  edge-strength map: max over 40 random cubic Bezier
 *                      curves of w * exp(-d^2 / (2 * 1.5^2)), w ~ U[0.3, 1],
 *                      d = distance to the curve, plus U[0, 0.05] noise,
 *                      clamped to [0, 1].
 *
 * Shape parameters come from one RNG stream; every image row draws its noise
 * from its own stream (seed mixed with the row index), so the result does not
 * depend on the number of OpenMP threads.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

typedef struct { uint64_t s; int have_spare; double spare; } rng_t;

static uint64_t splitmix_next(rng_t *r) {
    uint64_t z = (r->s += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

static void rng_init(rng_t *r, uint64_t seed) { r->s = seed; r->have_spare = 0; r->spare = 0.0; }

/* uniform in [0, 1) with 53 random bits */
static double u01(rng_t *r) { return (double)(splitmix_next(r) >> 11) * (1.0 / 9007199254740992.0); }

/* standard normal, Box-Muller (both values of each pair are used) */
static double normal01(rng_t *r) {
    if (r->have_spare) { r->have_spare = 0; return r->spare; }
    double u1 = 1.0 - u01(r);          /* (0, 1] */
    double u2 = u01(r);
    double rad = sqrt(-2.0 * log(u1));
    double ang = 6.283185307179586476925 * u2;
    r->spare = rad * sin(ang);
    r->have_spare = 1;
    return rad * cos(ang);
}

static uint64_t row_stream_seed(uint64_t seed, uint64_t row, uint64_t salt) {
    rng_t r; rng_init(&r, seed ^ (0xD1B54A32D192ED03ULL * (row + 1)) ^ salt);
    return splitmix_next(&r);
}

#define TOED_MAX_POLY 6

typedef struct {
    int is_disk;
    double cx, cy, r;                    /* disk */
    int nv;                              /* polygon */
    double vx[TOED_MAX_POLY], vy[TOED_MAX_POLY];
    double ymin, ymax;
    double gray;
} shape_t;

static int cmp_double(const void *a, const void *b) {
    double x = *(const double *)a, y = *(const double *)b;
    return (x > y) - (x < y);
}

/* x-interval (xl, xr) of the shape on the horizontal line y; 0 if empty */
static int shape_span(const shape_t *s, double y, double *xl, double *xr) {
    if (y <= s->ymin || y >= s->ymax) return 0;
    if (s->is_disk) {
        double dy = y - s->cy;
        double h2 = s->r * s->r - dy * dy;
        if (h2 <= 0.0) return 0;
        double h = sqrt(h2);
        *xl = s->cx - h; *xr = s->cx + h;
        return 1;
    }
    double lo = 1e300, hi = -1e300;
    int hits = 0;
    for (int k = 0; k < s->nv; ++k) {
        int k1 = (k + 1) % s->nv;
        double y0 = s->vy[k], y1 = s->vy[k1];
        if ((y0 <= y && y < y1) || (y1 <= y && y < y0)) {
            double t = (y - y0) / (y1 - y0);
            double x = s->vx[k] + t * (s->vx[k1] - s->vx[k]);
            if (x < lo) lo = x;
            if (x > hi) hi = x;
            ++hits;
        }
    }
    if (hits < 2 || !(hi > lo)) return 0;
    *xl = lo; *xr = hi;
    return 1;
}

int toed_scene_default_shapes(int H, int W) {
    (void)W;
    int k = H / 540; if (k < 1) k = 1;
    return 24 * k;
}

/* Returns 0 on success, -1 on bad arguments. */
int toed_scene_u8(uint64_t seed, int H, int W, int n_shapes, double size_scale, uint8_t *out) {
    if (H <= 0 || W <= 0 || out == NULL) return -1;
    int k = H / 540; if (k < 1) k = 1;
    if (n_shapes <= 0) n_shapes = 24 * k;
    if (size_scale <= 0.0) size_scale = (double)k;

    rng_t r; rng_init(&r, seed);
    double bg = 40.0 + 160.0 * u01(&r);
    shape_t *sh = (shape_t *)calloc((size_t)n_shapes, sizeof(shape_t));
    if (!sh) return -1;
    for (int s = 0; s < n_shapes; ++s) {
        shape_t *p = &sh[s];
        if ((s & 1) == 0) {
            p->is_disk = 1;
            p->cx = W * u01(&r);
            p->cy = H * u01(&r);
            p->r = (10.0 + 70.0 * u01(&r)) * size_scale;
            p->ymin = p->cy - p->r; p->ymax = p->cy + p->r;
        } else {
            p->is_disk = 0;
            double cx = W * u01(&r), cy = H * u01(&r);
            double bw = (20.0 + 130.0 * u01(&r)) * size_scale;
            double bh = (20.0 + 130.0 * u01(&r)) * size_scale;
            int nv = 3 + (int)floor(4.0 * u01(&r));
            if (nv > TOED_MAX_POLY) nv = TOED_MAX_POLY;
            double ang[TOED_MAX_POLY];
            for (int v = 0; v < nv; ++v) ang[v] = 6.283185307179586476925 * u01(&r);
            qsort(ang, (size_t)nv, sizeof(double), cmp_double);
            p->nv = nv;
            p->ymin = 1e300; p->ymax = -1e300;
            for (int v = 0; v < nv; ++v) {
                p->vx[v] = cx + 0.5 * bw * cos(ang[v]);
                p->vy[v] = cy + 0.5 * bh * sin(ang[v]);
                if (p->vy[v] < p->ymin) p->ymin = p->vy[v];
                if (p->vy[v] > p->ymax) p->ymax = p->vy[v];
            }
        }
        p->gray = 255.0 * u01(&r);
    }

    int fail = 0;
#pragma omp parallel
    {
        double *sub = (double *)malloc(sizeof(double) * 16 * (size_t)W);
        if (!sub) {
#pragma omp atomic write
            fail = 1;
        }
#pragma omp for schedule(dynamic, 4)
        for (int y = 0; y < H; ++y) {
            if (!sub) continue;
            const int SW = 4 * W;               /* subsamples per subsample row */
            for (int i = 0; i < 4 * SW; ++i) sub[i] = bg;
            for (int s = 0; s < n_shapes; ++s) {
                const shape_t *p = &sh[s];
                if (p->ymax <= y || p->ymin >= y + 1) continue;
                for (int kr = 0; kr < 4; ++kr) {
                    double ys = y + (kr + 0.5) * 0.25;
                    double xl, xr;
                    if (!shape_span(p, ys, &xl, &xr)) continue;
                    long mlo = (long)floor(4.0 * xl - 0.5) + 1;
                    long mhi = (long)ceil(4.0 * xr - 0.5) - 1;
                    if (mlo < 0) mlo = 0;
                    if (mhi > SW - 1) mhi = SW - 1;
                    double *row = sub + (size_t)kr * SW;
                    for (long m = mlo; m <= mhi; ++m) row[m] = p->gray;
                }
            }
            rng_t nr; rng_init(&nr, row_stream_seed(seed, (uint64_t)y, 0x5ce7e5ULL));
            for (int x = 0; x < W; ++x) {
                double acc = 0.0;
                for (int kr = 0; kr < 4; ++kr)
                    for (int kc = 0; kc < 4; ++kc) acc += sub[(size_t)kr * SW + 4 * x + kc];
                double v = acc * (1.0 / 16.0) + 2.0 * normal01(&nr);
                v = floor(v + 0.5);
                if (v < 0.0) v = 0.0;
                if (v > 255.0) v = 255.0;
                out[(size_t)y * W + x] = (uint8_t)v;
            }
        }
        free(sub);
    }
    free(sh);
    return fail ? -1 : 0;
}

/* ---------------------------------------------------------------------------
 * Bezier edge-strength map
 * ------------------------------------------------------------------------- */
typedef struct { float x0, y0, x1, y1, w; } seg_t;

int toed_bezier_map_f32(uint64_t seed, int H, int W, int n_curves, float *out) {
    if (H <= 0 || W <= 0 || out == NULL) return -1;
    if (n_curves <= 0) n_curves = 40;
    const double sig = 1.5, inv2s2 = 1.0 / (2.0 * sig * sig), reach = 9.0;

    rng_t r; rng_init(&r, seed);
    size_t cap = 1024, n = 0;
    seg_t *seg = (seg_t *)malloc(cap * sizeof(seg_t));
    if (!seg) return -1;
    for (int c = 0; c < n_curves; ++c) {
        double px[4], py[4];
        for (int k = 0; k < 4; ++k) { px[k] = W * u01(&r); py[k] = H * u01(&r); }
        double w = 0.3 + 0.7 * u01(&r);
        double len = 0.0;
        for (int k = 0; k < 3; ++k) len += hypot(px[k + 1] - px[k], py[k + 1] - py[k]);
        int ns = (int)ceil(len / 0.25); if (ns < 1) ns = 1;
        double qx = px[0], qy = py[0];
        for (int i = 1; i <= ns; ++i) {
            double t = (double)i / ns, u = 1.0 - t;
            double b0 = u * u * u, b1 = 3 * u * u * t, b2 = 3 * u * t * t, b3 = t * t * t;
            double x = b0 * px[0] + b1 * px[1] + b2 * px[2] + b3 * px[3];
            double y = b0 * py[0] + b1 * py[1] + b2 * py[2] + b3 * py[3];
            if (n == cap) {
                cap *= 2;
                seg_t *ns2 = (seg_t *)realloc(seg, cap * sizeof(seg_t));
                if (!ns2) { free(seg); return -1; }
                seg = ns2;
            }
            seg[n].x0 = (float)qx; seg[n].y0 = (float)qy;
            seg[n].x1 = (float)x;  seg[n].y1 = (float)y;
            seg[n].w = (float)w;
            ++n;
            qx = x; qy = y;
        }
    }

    /* bin segments into 32-row bands by their reach-expanded y-extent */
    const int BAND = 32;
    int nb = (H + BAND - 1) / BAND;
    size_t *bcount = (size_t *)calloc((size_t)nb + 1, sizeof(size_t));
    if (!bcount) { free(seg); return -1; }
    for (size_t i = 0; i < n; ++i) {
        double y0 = fmin(seg[i].y0, seg[i].y1) - reach, y1 = fmax(seg[i].y0, seg[i].y1) + reach;
        int b0 = (int)floor(y0 / BAND), b1 = (int)floor(y1 / BAND);
        if (b0 < 0) b0 = 0;
        if (b1 > nb - 1) b1 = nb - 1;
        for (int b = b0; b <= b1; ++b) bcount[b + 1]++;
    }
    for (int b = 0; b < nb; ++b) bcount[b + 1] += bcount[b];
    size_t *bidx = (size_t *)malloc((bcount[nb] + 1) * sizeof(size_t));
    size_t *fillp = (size_t *)malloc((size_t)nb * sizeof(size_t));
    if (!bidx || !fillp) { free(seg); free(bcount); free(bidx); free(fillp); return -1; }
    for (int b = 0; b < nb; ++b) fillp[b] = bcount[b];
    for (size_t i = 0; i < n; ++i) {
        double y0 = fmin(seg[i].y0, seg[i].y1) - reach, y1 = fmax(seg[i].y0, seg[i].y1) + reach;
        int b0 = (int)floor(y0 / BAND), b1 = (int)floor(y1 / BAND);
        if (b0 < 0) b0 = 0;
        if (b1 > nb - 1) b1 = nb - 1;
        for (int b = b0; b <= b1; ++b) bidx[fillp[b]++] = i;
    }

#pragma omp parallel for schedule(dynamic, 1)
    for (int b = 0; b < nb; ++b) {
        int ylo = b * BAND, yhi = ylo + BAND; if (yhi > H) yhi = H;
        float *blk = out + (size_t)ylo * W;
        for (size_t i = 0; i < (size_t)(yhi - ylo) * W; ++i) blk[i] = 0.0f;
        for (size_t e = bcount[b]; e < bcount[b + 1]; ++e) {
            const seg_t *s = &seg[bidx[e]];
            double ax = s->x0, ay = s->y0, bx = s->x1, by = s->y1;
            double dx = bx - ax, dy = by - ay, l2 = dx * dx + dy * dy;
            int x0 = (int)floor(fmin(ax, bx) - reach), x1 = (int)ceil(fmax(ax, bx) + reach);
            int y0 = (int)floor(fmin(ay, by) - reach), y1 = (int)ceil(fmax(ay, by) + reach);
            if (x0 < 0) x0 = 0;
            if (x1 > W - 1) x1 = W - 1;
            if (y0 < ylo) y0 = ylo;
            if (y1 > yhi - 1) y1 = yhi - 1;
            for (int y = y0; y <= y1; ++y) {
                for (int x = x0; x <= x1; ++x) {
                    /* distance from the pixel centre to the segment */
                    double qx = x - ax, qy = y - ay;
                    double t = l2 > 0 ? (qx * dx + qy * dy) / l2 : 0.0;
                    if (t < 0) t = 0;
                    if (t > 1) t = 1;
                    double ex = qx - t * dx, ey = qy - t * dy;
                    double d2 = ex * ex + ey * ey;
                    if (d2 > reach * reach) continue;
                    float v = (float)(s->w * exp(-d2 * inv2s2));
                    float *o = &out[(size_t)y * W + x];
                    if (v > *o) *o = v;
                }
            }
        }
        for (int y = ylo; y < yhi; ++y) {
            rng_t nr; rng_init(&nr, row_stream_seed(seed, (uint64_t)y, 0xbe21e7ULL));
            for (int x = 0; x < W; ++x) {
                float v = out[(size_t)y * W + x] + (float)(0.05 * u01(&nr));
                if (v < 0.0f) v = 0.0f;
                if (v > 1.0f) v = 1.0f;
                out[(size_t)y * W + x] = v;
            }
        }
    }
    free(seg); free(bcount); free(bidx); free(fillp);
    return 0;
}

/* Fixed per-seed loss weights (a, b, c, d) ~ N(0, 1) for the edgel loss
 * sum(a*x + b*y + c*theta + d*strength). */
void toed_loss_weights(uint64_t seed, double w[4]) {
    rng_t r; rng_init(&r, seed ^ 0x10557e1ULL);
    for (int k = 0; k < 4; ++k) w[k] = normal01(&r);
}
