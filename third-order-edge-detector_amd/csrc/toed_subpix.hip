// Subpixel third-order correction of dense edge maps on MI355X (gfx950):
// the GPU form of the reference's SubpixelTOCorrectionCPU<T>
// (subpix_TO_correction/subpix_to_correction.cpp).
//
// Per map (any number of maps of any sizes in one plan, one sigma):
//   R  sp_resize_kernel       interp_n = n > 0 only: the map (and a given
//        orientation) bilinearly resized to the work grid
//        ((H - 1) 2^n + 1) x ((W - 1) 2^n + 1) (:69-90, :226-240); every step
//        below then runs on the work grid (NMS margin 3 * 2^n, list x, y / 2^n,
//        isolation radius 2^n r, TO maps gathered onto the H x W grid).
//   A  sp_orient_nms_kernel   one workgroup per 16 x 32 tile.  Orientation:
//        given, or estimated from the map (:307-375): triangle smoothing r = 4
//        with half-sample-symmetric borders (row pass, column pass, FP64, same
//        summation order), central differences (one-sided at the map border),
//        theta = mod(atan(Oyy sgn(-Oxy) / (Oxx + 1e-5)), pi), direction
//        (cos, sin)(theta) (:296-305).  Writes the planes Hx = E dirx,
//        Hy = E diry (FP32) and runs the octant NMS with the parabola in FP64
//        (:413-519); survivors write their FP64 position and strength into
//        dense per-pixel planes and set a bit in a per-row 32-bit mask word.
//   B1 sp_isolate_kernel      isolated-edgel removal (:558-621): a survivor is
//        kept when another survivor lies within the radius; candidates are the
//        survivors within floor(r + 2 sqrt 2) pixels (edgels lie within
//        sqrt 2 of their pixel) -> kept mask words.
//   C  sp_scan_chunks / sp_scan_maps / sp_prefix_kernel   exclusive scan of
//        the kept-mask popcounts in raster order (chunks of 1024 words, then
//        per map): the count, and the rank of each word's first kept edgel.
//   B2 sp_correct_kernel<R>   one workgroup per tile with kept edgels: the
//        eight circular separable Gaussian-derivative convolutions
//        (:377-411, :641-648) of the tile plus a 3-pixel halo in LDS (FP32,
//        x pass then y pass), Keys cubic sampling of the ten planes at every
//        kept edgel (FP64, :521-556), the third-order orientation (:663-700,
//        FP64); each edgel is written as row (rank) of the N x 4 list
//        [x, y, orientation, strength] (:702-729).
//   E  sp_maps_kernel         optional TO_edgemap / TO_orientation: each pixel
//        gathers the kept edgels of its 3 x 3 neighbourhood whose rounded
//        position is the pixel and takes the last in list order (:708-715).
// No MFMA: short separable stencils.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <vector>

#include "toed_capi.h"
#include "toed_dpp.hpp"

namespace toedsp {

constexpr int TH = 16;          // tile rows
constexpr int TW = 32;          // tile columns = one mask word per tile row
constexpr int NT = 256;         // threads per workgroup
constexpr int MAXR = 12;        // largest filter radius ceil(4 sigma)
constexpr int MAXK = 8;         // isolation reach in work pixels of the default isolation kernel
constexpr int MAXK_WIDE = 12;   // ... and of the wide one (radius 2^n r on an interp_n = n work grid)
constexpr int CHUNKW = 1024;    // mask words per scan chunk
constexpr int HALO = 3;         // cubic stencil reach around an edgel pixel

struct SpTaps {
    float k[3][2 * MAXR + 1];   // G, G', G'' at integer offsets -R..R (:194-214)
    double r2;                  // isolation radius squared
    int R, K, keep_all, pad;    // K = isolation reach; keep_all when radius <= 0
};

struct SpJob {
    const float *map;
    const float *orient;        // null: estimated
    float *hx, *hy;             // H x W planes
    double2 *pos;               // H x W: FP64 subpixel position of survivors
    float2 *val;                // H x W: strength, orientation of survivors
    uint32_t *vmask, *kmask;    // H x nw survivor / kept words
    int *nsurv;                 // survivors of the whole map, counted up to 2 (own 256-byte slot)
    float *edges;
    int *index;
    float *to_e, *to_o;
    int *count_out, *flags_out;
    long long capacity;
    double thresh;
    int H, W, pitch, opitch;
    int nw, ntx, nty, tile_begin;
    int chunk_begin, nchunks, tpitch, pad;
    uint32_t *wpre;             // H x nw: rank of each kept word's first edgel
    uint8_t *tflag;             // one byte per tile (raster order): nonzero when the tile has kept edgels
    int *tlist;                 // those tiles as a list: their number, then their indices
    // differentiable form (null unless enabled on the plan)
    float4 *tq;                 // H x W: dirx, diry, dOxx/dtheta, dOyy/dtheta coefficients
    double2 *dirs;              // H x W: direction of survivors
    float4 *rec;                // 8 per edgel (by rank): the ten cubic sums, their x and y derivatives,
                                // the edgel's map pixel i W + j (int bits of float 30)
    float4 *brec;               // 4 per edgel (by rank): the ten sample adjoints, the cubic
                                // stencil's fraction (fx, fy) and base (int bits of xf, yf)
    int *gcount;                // kept edgels of the map (device int)
    float *gth;                 // H x W scratch: dL/dtheta (cleared, then the edgels' own terms and
                                // the tile kernel's E (dirx dL/dHy - diry dL/dHx))
    // work grid (interp_n = slog2 > 0, :92-105, :226-240): H x W above is the
    // work grid ((IH - 1) 2^n + 1) x ((IW - 1) 2^n + 1); map / orient point at
    // the plan's work planes, resized from the caller's IH x IW arrays src_map /
    // src_orient (rows of spitch / sopitch); the TO maps stay IH x IW.  With
    // slog2 = 0, IH = H and IW = W.
    const float *src_map, *src_orient;
    int IH, IW, spitch, sopitch;
    int margin, slog2;          // NMS margin 3 * 2^n (:627-629)
};

// upstream gradient and output of one map's backward pass
struct SpGrad {
    const float4 *gout;         // dL/d[x, y, orientation, strength], raster order
    float *gmap;                // dL/dE, accumulated, rows of gpitch
    int gpitch, pad;
};

#define SP_CONST __attribute__((address_space(4)))

// Grids are (tiles or chunks of the largest map, maps): the map of a
// workgroup is blockIdx.y, read with scalar loads.
__device__ __forceinline__ SpJob load_job(const SpJob *jobs) {
#if defined(__HIP_DEVICE_COMPILE__)
    return ((const SP_CONST SpJob *)jobs)[blockIdx.y];
#else
    return jobs[0];
#endif
}

// The map's gradient pointers, read like the job (scalar loads; accesses
// through them compile to global, not flat, instructions).
__device__ __forceinline__ SpGrad load_grad(const SpGrad *grads) {
#if defined(__HIP_DEVICE_COMPILE__)
    return ((const SP_CONST SpGrad *)grads)[blockIdx.y];
#else
    return grads[0];
#endif
}

// half-sample-symmetric index (:47-55) and periodic index (:40-45)
__device__ __forceinline__ int wrap_sym(int i, int n) {
    if (n <= 1) return 0;
    while (i < 0 || i >= n) i = (i < 0) ? -i - 1 : 2 * n - i - 1;
    return i;
}
__device__ __forceinline__ int wrap_circ(int i, int n) {
    int r = i % n;
    return r < 0 ? r + n : r;
}

__device__ __forceinline__ double mod_pi(double a) {
    double r = fmod(a, M_PI);
    if (r < 0) r += M_PI;
    return r;
}

// direction (cos, sin)(mod(atan(t), pi)) without trigonometry: atan(t) lies
// in (-pi/2, pi/2); a negative angle moves up by pi, which flips both signs.
__device__ __forceinline__ void dir_from_slope(double t, double &dx, double &dy) {
    if (fabs(t) <= 1e150 && !(t < 0.0 && t > -1e-6)) {
        const double inv = 1.0 / sqrt(1.0 + t * t);
        if (t < 0.0) {
            dx = -inv;
            dy = -t * inv;
        } else {
            dx = inv;
            dy = t * inv;
        }
        return;
    }
    // infinities, NaN, and small negative slopes, where atan(t) + pi may round
    // to pi and the angle wraps to 0: the trigonometric route of the reference
    const double a = mod_pi(mod_pi(atan(t)));
    dx = cos(a);
    dy = sin(a);
}

__device__ __forceinline__ double tri_w(int q) { return (double)(5 - (q < 0 ? -q : q)) / 25.0; }

// ---------------------------------------------------------------------------
// A: orientation, H planes, NMS
constexpr int AE_R = TH + 12, AE_C = TW + 12;   // map rows/cols, tile +- 6
constexpr int AT_C = TW + 4;                    // row-pass columns, tile +- 2
constexpr int AS_R = TH + 4;                    // smoothed rows, tile +- 2

__global__ __launch_bounds__(NT) void sp_orient_nms_kernel(const SpJob *jobs, int njobs) {
    __shared__ float Es[AE_R][AE_C];
    __shared__ double Tr[AE_R][AT_C];
    __shared__ double Ss[AS_R][AT_C];
    const SpJob J = load_job(jobs);
    const int lt = blockIdx.x;
    if (lt >= J.ntx * J.nty) return;
    const int ty = lt / J.ntx, tx = lt - ty * J.ntx;
    const int y0 = ty * TH, x0 = tx * TW;
    const int H = J.H, W = J.W;
    const bool est = (J.orient == nullptr);
    // no pixel of the tile is within two pixels of the map border
    const bool inner2 = x0 >= 2 && x0 + TW <= W - 2 && y0 >= 2 && y0 + TH <= H - 2;

    if (y0 >= 6 && y0 + TH + 6 <= H && x0 >= 6 && x0 + TW + 6 <= W) {
        // interior tile: direct indexing, no border folding
        const float *src = J.map + (long long)(y0 - 6) * J.pitch + (x0 - 6);
        for (int idx = threadIdx.x; idx < AE_R * AE_C; idx += NT) {
            const int r = idx / AE_C, c = idx - r * AE_C;
            Es[r][c] = src[(long long)r * J.pitch + c];
        }
    } else {
        for (int idx = threadIdx.x; idx < AE_R * AE_C; idx += NT) {
            const int r = idx / AE_C, c = idx - r * AE_C;
            const int gy = wrap_sym(y0 - 6 + r, H), gx = wrap_sym(x0 - 6 + c, W);
            Es[r][c] = J.map[(long long)gy * J.pitch + gx];
        }
    }
    __syncthreads();
    if (est) {
        // row pass (:329-336): tmp(i, j) = sum_q E(i, sym(j - q)) tri[q], four
        // outputs per thread from one 12-wide window (same order of terms per output)
        constexpr int NO = 4, NCB = AT_C / NO, NRB = AS_R / NO;
        static_assert(AT_C % NO == 0 && AS_R % NO == 0, "whole output blocks");
        for (int idx = threadIdx.x; idx < AE_R * NCB; idx += NT) {
            const int r = idx / NCB, c0 = (idx - r * NCB) * NO;
            double w[NO + 8];
#pragma unroll
            for (int t = 0; t < NO + 8; ++t) w[t] = (double)Es[r][c0 + t];
#pragma unroll
            for (int u = 0; u < NO; ++u) {
                double acc = 0.0;
#pragma unroll
                for (int q = -4; q <= 4; ++q) acc += w[u + 4 - q] * tri_w(q);
                Tr[r][c0 + u] = acc;
            }
        }
        __syncthreads();
        // column pass (:337-344), four outputs per thread down a column
        for (int idx = threadIdx.x; idx < NRB * AT_C; idx += NT) {
            const int rb = idx / AT_C, c = idx - rb * AT_C, r0 = rb * NO;
            double w[NO + 8];
#pragma unroll
            for (int t = 0; t < NO + 8; ++t) w[t] = Tr[r0 + t][c];
#pragma unroll
            for (int u = 0; u < NO; ++u) {
                double acc = 0.0;
#pragma unroll
                for (int p = -4; p <= 4; ++p) acc += w[u + 4 - p] * tri_w(p);
                Ss[r0 + u][c] = acc;
            }
        }
        __syncthreads();
    }

    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    // the thread's two pixels unrolled: two independent FP64 chains in flight
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        const int p = threadIdx.x + k * NT;
        const int li = p / TW, lj = p - li * TW;
        const int gi = y0 + li, gj = x0 + lj;
        bool surv = false;
        if (gi < H && gj < W) {
            double dx, dy;
            if (est) {
                // smoothed map S(i, j) for in-map (i, j) of the tile +- 2 window
                auto S = [&](int i, int j) -> double { return Ss[i - y0 + 2][j - x0 + 2]; };
                // gradient2 (:346-357): central differences, one-sided at the border
                auto Ox = [&](int i, int j) -> double {
                    if (j == 0) return S(i, 1) - S(i, 0);
                    if (j == W - 1) return S(i, W - 1) - S(i, W - 2);
                    return 0.5 * (S(i, j + 1) - S(i, j - 1));
                };
                auto Oy = [&](int i, int j) -> double {
                    if (i == 0) return S(1, j) - S(0, j);
                    if (i == H - 1) return S(H - 1, j) - S(H - 2, j);
                    return 0.5 * (S(i + 1, j) - S(i - 1, j));
                };
                double oxx, oxy, oyy;
                if (inner2) {
                    // central differences of central differences: the general branch's
                    // operations for a pixel two or more pixels inside the map
                    oxx = 0.5 * (0.5 * (S(gi, gj + 2) - S(gi, gj)) - 0.5 * (S(gi, gj) - S(gi, gj - 2)));
                    oxy = 0.5 * (0.5 * (S(gi + 1, gj + 1) - S(gi - 1, gj + 1)) - 0.5 * (S(gi + 1, gj - 1) - S(gi - 1, gj - 1)));
                    oyy = 0.5 * (0.5 * (S(gi + 2, gj) - S(gi, gj)) - 0.5 * (S(gi, gj) - S(gi - 2, gj)));
                } else {
                if (gj == 0) { oxx = Ox(gi, 1) - Ox(gi, 0); oxy = Oy(gi, 1) - Oy(gi, 0); }
                else if (gj == W - 1) { oxx = Ox(gi, W - 1) - Ox(gi, W - 2); oxy = Oy(gi, W - 1) - Oy(gi, W - 2); }
                else { oxx = 0.5 * (Ox(gi, gj + 1) - Ox(gi, gj - 1)); oxy = 0.5 * (Oy(gi, gj + 1) - Oy(gi, gj - 1)); }
                if (gi == 0) oyy = Oy(1, gj) - Oy(0, gj);
                else if (gi == H - 1) oyy = Oy(H - 1, gj) - Oy(H - 2, gj);
                else oyy = 0.5 * (Oy(gi + 1, gj) - Oy(gi - 1, gj));
                }
                const double sgn = (oxy < 0) ? 1.0 : ((oxy > 0) ? -1.0 : 0.0);
                const double a = oxx + 1e-5, t = oyy * sgn / a;
                dir_from_slope(t, dx, dy);
                if (J.tq) {
                    // theta = atan(t) (+ pi): dtheta/dOxx = -t / ((1 + t^2) a), dtheta/dOyy = sgn / ((1 + t^2) a)
                    const double c = 1.0 / ((1.0 + t * t) * a);
                    const double cxx = -t * c, cyy = sgn * c;
                    J.tq[(long long)gi * W + gj] = make_float4((float)dx, (float)dy, isfinite(cxx) ? (float)cxx : 0.f,
                                                               isfinite(cyy) ? (float)cyy : 0.f);
                }
            } else {
                const double a = mod_pi((double)J.orient[(long long)gi * J.opitch + gj]);
                dx = cos(a);
                dy = sin(a);
                if (J.tq) J.tq[(long long)gi * W + gj] = make_float4((float)dx, (float)dy, 0.f, 0.f);
            }
            const double e = (double)Es[li + 6][lj + 6];
            const long long o = (long long)gi * W + gj;
            J.hx[o] = (float)(e * dx);
            J.hy[o] = (float)(e * dy);
            // NMS (:446-516): rows/cols [margin + 1, n - margin - 2), margin 3 * 2^n
            if (gi > J.margin && gi < H - J.margin - 2 && gj > J.margin && gj < W - J.margin - 2 && !(e <= J.thresh) &&
                !(fabs(dx) < 1e-5 && fabs(dy) < 1e-5)) {
                auto M = [&](int di, int dj) -> double { return (double)Es[li + 6 + di][lj + 6 + dj]; };
                const double md = sqrt(dx * dx + dy * dy);
                const double nx = dx / md, ny = dy / md;
                double slope, fp, fm;
                if (dx >= 0 && dy >= 0) {
                    if (dx >= dy) { slope = ny / nx; fp = M(0, 1) * (1 - slope) + M(1, 1) * slope; fm = M(0, -1) * (1 - slope) + M(-1, -1) * slope; }
                    else { slope = nx / ny; fp = M(1, 0) * (1 - slope) + M(1, 1) * slope; fm = M(-1, 0) * (1 - slope) + M(-1, -1) * slope; }
                } else if (dx < 0 && dy >= 0) {
                    if (fabs(dx) < dy) { slope = -nx / ny; fp = M(1, 0) * (1 - slope) + M(1, -1) * slope; fm = M(-1, 0) * (1 - slope) + M(-1, 1) * slope; }
                    else { slope = -ny / nx; fp = M(0, -1) * (1 - slope) + M(1, -1) * slope; fm = M(0, 1) * (1 - slope) + M(-1, 1) * slope; }
                } else if (dx < 0 && dy < 0) {
                    if (fabs(dx) >= fabs(dy)) { slope = ny / nx; fp = M(0, -1) * (1 - slope) + M(-1, -1) * slope; fm = M(0, 1) * (1 - slope) + M(1, 1) * slope; }
                    else { slope = nx / ny; fp = M(-1, 0) * (1 - slope) + M(-1, -1) * slope; fm = M(1, 0) * (1 - slope) + M(1, 1) * slope; }
                } else {
                    if (dx < fabs(dy)) { slope = -nx / ny; fp = M(-1, 0) * (1 - slope) + M(-1, 1) * slope; fm = M(1, 0) * (1 - slope) + M(1, -1) * slope; }
                    else { slope = -ny / nx; fp = M(0, 1) * (1 - slope) + M(-1, 1) * slope; fm = M(0, -1) * (1 - slope) + M(1, -1) * slope; }
                }
                const double f = e;
                if ((f > fm && f > fp) || (f > fm && f >= fp) || (f >= fm && f > fp)) {
                    const double s = sqrt(1 + slope * slope);
                    const double A = (fm + fp - 2 * f) / (2 * s * s);
                    const double B = (fp - fm) / (2 * s);
                    if (!(fabs(A) < 1e-16)) {
                        const double ss = -B / (2 * A);
                        const double mf = A * ss * ss + B * ss + f;
                        if (fabs(ss) <= sqrt(2.0)) {
                            J.pos[o] = make_double2(gj + ss * nx, gi + ss * ny);
                            if (J.dirs) J.dirs[o] = make_double2(dx, dy);
                            J.val[o] = make_float2((float)fabs(mf), atan2f((float)dx, (float)-dy));   // unit direction: FP32 angle
                            surv = true;
                        }
                    }
                }
            }
        }
        const unsigned long long b = __ballot(surv);
        if ((lane & 31) == 0) {
            const int row = y0 + 2 * wave + 8 * k + (lane >> 5);
            if (row < H) J.vmask[(long long)row * J.nw + tx] = (uint32_t)(b >> lane);
        }
        // only "fewer than two survivors" matters (:563-564): stop counting at 2, so the
        // map's counter line sees a handful of atomics instead of one per wavefront
        if (lane == 0 && b && __hip_atomic_load(J.nsurv, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < 2)
            atomicAdd(J.nsurv, (int)__popcll(b));
    }
}

// ---------------------------------------------------------------------------
// B: isolation, circular convolutions, cubic sampling, third-order orientation
__device__ __forceinline__ double keys(double s) {
    const double a = -0.5;
    const double x = fabs(s);
    if (x <= 1.0) return ((a + 2) * x - (a + 3)) * x * x + 1;
    if (x < 2.0) return ((a * x - 5 * a) * x + 8 * a) * x - 4 * a;
    return 0.0;
}

// derivative of the Keys kernel
__device__ __forceinline__ double keys_d(double s) {
    const double a = -0.5;
    const double x = fabs(s), sg = s < 0 ? -1.0 : 1.0;
    if (x <= 1.0) return sg * ((3 * (a + 2) * x - 2 * (a + 3)) * x);
    if (x < 2.0) return sg * ((3 * a * x - 10 * a) * x + 8 * a);
    return 0.0;
}

template <int R> struct CGeo {
    static constexpr int NK = 2 * R + 1;
    static constexpr int PR = TH + 2 * HALO, PC = TW + 2 * HALO;   // plane region
    static constexpr int IR = PR + 2 * R, IC = PC + 2 * R;         // input region
    static constexpr int IN = 2 * IR * IC;                         // E, or Hx and Hy
    static constexpr int XN = IR * PC;                             // x-pass, one plane
    static constexpr int YN = 8 * PR * PC;                         // 8 planes
    static constexpr int FLOATS = IN + XN + YN;
};

// The plane rows r0 .. r1 and columns c0 .. c1 (even .. odd) a tile's cubic
// stencils can reach: a kept edgel at tile pixel (i, j) lies within sqrt 2
// of it, so its 4 x 4 stencil stays in plane rows i .. i + 2 HALO and columns
// j .. j + 2 HALO.  In is needed over rows r0 .. r1 + 2R and columns
// c0 .. c1 + 2R, the x pass there, the y pass over the box itself.
struct PlaneBox {
    int r0, r1, c0, c1;
};

// E (WHICH = 0, into slot 0) or Hx, Hy (WHICH = 1, into slots 0 and 1) over
// the plane region + R (periodic borders) into In, rows and columns of the
// box only
template <int R, int WHICH>
__device__ __forceinline__ void load_tile_in(const SpJob &J, float *In, int y0, int x0, const PlaneBox &bx) {
    using G = CGeo<R>;
    constexpr int IR = G::IR, IC = G::IC;
    const int H = J.H, W = J.W;
    const int iy0 = y0 - HALO - R, ix0 = x0 - HALO - R;
    const int lr = bx.r1 + 2 * R - bx.r0 + 1, lc = bx.c1 + 2 * R - bx.c0 + 1, ln = lr * lc;
    if (iy0 >= 0 && iy0 + IR <= H && ix0 >= 0 && ix0 + IC <= W) {
        constexpr int U = 4;
        for (int base = 0; base < ln; base += U * NT) {
            float v[2][U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int k = base + threadIdx.x + u * NT;
                const int r = bx.r0 + k / lc, c = bx.c0 + k % lc;
                const bool ok = k < ln;
                const long long o = (long long)(iy0 + r) * W + (ix0 + c);
                if (WHICH == 0) {
                    v[0][u] = ok ? J.map[(long long)(iy0 + r) * J.pitch + (ix0 + c)] : 0.f;
                } else {
                    v[0][u] = ok ? J.hx[o] : 0.f;
                    v[1][u] = ok ? J.hy[o] : 0.f;
                }
            }
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int k = base + threadIdx.x + u * NT;
                if (k < ln) {
                    const int idx = (bx.r0 + k / lc) * IC + bx.c0 + k % lc;
                    In[idx] = v[0][u];
                    if (WHICH == 1) In[IR * IC + idx] = v[1][u];
                }
            }
        }
    } else {
        for (int k = threadIdx.x; k < ln; k += NT) {
            const int r = bx.r0 + k / lc, c = bx.c0 + k % lc, idx = r * IC + c;
            const int gy = wrap_circ(iy0 + r, H), gx = wrap_circ(ix0 + c, W);
            const long long o = (long long)gy * W + gx;
            if (WHICH == 0) {
                In[idx] = J.map[(long long)gy * J.pitch + gx];
            } else {
                In[idx] = J.hx[o];
                In[IR * IC + idx] = J.hy[o];
            }
        }
    }
}

// The circular Gaussian-derivative planes of one group for the tile + HALO,
// one x-pass plane at a time (x pass into X, y pass into Y); ends with a
// barrier.  Group 0 from E (In slot 0): E * G' -> Px (G), Pxy (G'); E * G ->
// Py (G'), Pyy (G''); E * G'' -> Pxx (G).  Group 1 from Hx, Hy (slots 0, 1):
// Hx * G' -> Hxx (G); Hy * G -> Hyy (G'); Hy * G' -> Hxy (G).  Planes Y: 0 Px,
// 1 Py, 2 Pxx, 3 Pyy, 4 Pxy, 5 Hxx, 6 Hxy, 7 Hyy.
// Two outputs per thread (adjacent columns in the x pass, adjacent rows in
// the y pass) share 2R + 2 loaded values.  Output u of a pass sums
// window[m + u] * k[2R - m], m = 0..2R (true convolution).
template <int R, int GRP>
__device__ __forceinline__ void conv_group(const float *In, float *X, float *Y, const SpTaps &tp, const PlaneBox &bx) {
    using G = CGeo<R>;
    constexpr int PR = G::PR, PC = G::PC, IR = G::IR, IC = G::IC;
    static_assert(PC % 2 == 0 && PR % 2 == 0, "two outputs per thread");
    const int xr = bx.r1 + 2 * R - bx.r0 + 1, ncp = (bx.c1 - bx.c0 + 1) / 2, nrp = (bx.r1 - bx.r0 + 1) / 2;
    // per x-pass plane: source slot, x filter, then up to two (Y plane, y filter)
    constexpr int SRC[2][3] = {{0, 0, 0}, {0, 1, 1}};
    constexpr int FX[2][3] = {{1, 0, 2}, {1, 0, 1}};
    constexpr int YA[2][3] = {{0, 1, 2}, {5, 7, 6}};
    constexpr int FA[2][3] = {{0, 1, 0}, {0, 1, 0}};
    constexpr int YB[2][3] = {{4, 3, -1}, {-1, -1, -1}};
    constexpr int FB[2][3] = {{1, 2, 0}, {0, 0, 0}};
#pragma unroll
    for (int pl = 0; pl < 3; ++pl) {
        const int fx = FX[GRP][pl];
        for (int idx = threadIdx.x; idx < xr * ncp; idx += NT) {
            const int r = bx.r0 + idx / ncp, c = bx.c0 + 2 * (idx % ncp);
            const float *src = In + SRC[GRP][pl] * IR * IC + r * IC + c;
            float a[2] = {0.f, 0.f};
#pragma unroll
            for (int m = 0; m <= 2 * R + 1; ++m) {
                const float v = src[m];
#pragma unroll
                for (int u = 0; u < 2; ++u) {
                    const int t = 2 * R - (m - u);
                    if (t < 0 || t > 2 * R) continue;
                    a[u] = fmaf(v, tp.k[fx][t], a[u]);
                }
            }
            X[r * PC + c] = a[0];
            X[r * PC + c + 1] = a[1];
        }
        __syncthreads();
        const int ya = YA[GRP][pl], fa = FA[GRP][pl], yb = YB[GRP][pl], fb = FB[GRP][pl];
        for (int idx = threadIdx.x; idx < nrp * (2 * ncp); idx += NT) {
            const int r = bx.r0 + 2 * (idx / (2 * ncp)), c = bx.c0 + idx % (2 * ncp);
            const float *xp = X + r * PC + c;
            float pa[2] = {0.f, 0.f}, pb[2] = {0.f, 0.f};
#pragma unroll
            for (int m = 0; m <= 2 * R + 1; ++m) {
                const float v = xp[m * PC];
#pragma unroll
                for (int u = 0; u < 2; ++u) {
                    const int t = 2 * R - (m - u);
                    if (t < 0 || t > 2 * R) continue;
                    pa[u] = fmaf(v, tp.k[fa][t], pa[u]);
                    if (yb >= 0) pb[u] = fmaf(v, tp.k[fb][t], pb[u]);
                }
            }
#pragma unroll
            for (int u = 0; u < 2; ++u) {
                const int o = (r + u) * PC + c;
                Y[ya * PR * PC + o] = pa[u];
                if (yb >= 0) Y[yb * PR * PC + o] = pb[u];
            }
        }
        __syncthreads();
    }
}

// B1: isolated-edgel removal (:558-621) for every tile: an edgel is kept
// when another survivor lies within the radius; survivors within
// floor(r + 2 sqrt 2) pixels are candidates (an edgel lies within sqrt 2 of
// its pixel).  Small LDS, so many tiles are in flight to hide the loads.
// KM: the largest reach K the instantiation holds (LDS is sized for it).
template <int KM>
__global__ __launch_bounds__(NT) void sp_isolate_kernel(const SpJob *jobs, SpTaps tp) {
    static_assert(2 * KM + 1 < 32 && KM <= TW, "one mask word each side");
    __shared__ uint32_t vm[TH + 2 * KM][3];
    __shared__ double2 posw[(TH + 2 * KM) * (TW + 2 * KM)];
    __shared__ int s_flag;

    const SpJob J = load_job(jobs);
    const int lt = blockIdx.x;
    if (lt >= J.ntx * J.nty) return;
    const int ty = lt / J.ntx, tx = lt - ty * J.ntx;
    const int y0 = ty * TH, x0 = tx * TW;
    const int H = J.H, W = J.W, K = tp.K;
    const int nsurv = *J.nsurv;

    // survivor words of rows y0-K .. y0+TH+K, word columns tx-1 .. tx+1
    if (threadIdx.x == 0) s_flag = 0;
    for (int idx = threadIdx.x; idx < (TH + 2 * K) * 3; idx += NT) {
        const int r = idx / 3, c = idx - r * 3;
        const int gy = y0 - K + r, wc = tx - 1 + c;
        vm[r][c] = (gy >= 0 && gy < H && wc >= 0 && wc < J.nw) ? J.vmask[(long long)gy * J.nw + wc] : 0u;
    }
    __syncthreads();
    if (threadIdx.x < TH && vm[K + threadIdx.x][1]) s_flag = 1;
    __syncthreads();
    const bool any_surv = s_flag != 0;
    __syncthreads();
    const bool keep_all = tp.keep_all || (nsurv < 2);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;

    // the survivors' positions of the tile + K window are staged in LDS
    const int WR = TH + 2 * K, WC = TW + 2 * K;
    auto survivor = [&](int r, int c) -> bool {      // window (r, c): tile (r - K, c - K)
        const int cc = TW - K + c;                   // column in the 3-word mask window
        return (vm[r][cc >> 5] >> (cc & 31)) & 1u;
    };
    if (any_surv && !keep_all) {
        // all of a thread's loads are issued before its LDS stores
        constexpr int UW = ((TH + 2 * KM) * (TW + 2 * KM) + NT - 1) / NT;
        double2 pv[UW];
        bool sv[UW];
#pragma unroll
        for (int u = 0; u < UW; ++u) {
            const int idx = threadIdx.x + u * NT;
            const int r = idx / WC, c = idx - r * WC;
            sv[u] = idx < WR * WC && survivor(r, c);
            pv[u] = sv[u] ? J.pos[(long long)(y0 - K + r) * W + (x0 - K + c)] : make_double2(0.0, 0.0);
        }
#pragma unroll
        for (int u = 0; u < UW; ++u)
            if (sv[u]) posw[threadIdx.x + u * NT] = pv[u];
    }
    __syncthreads();
#pragma unroll 1
    for (int k = 0; k < 2; ++k) {
        const int p = threadIdx.x + k * NT;
        const int li = p / TW, lj = p - li * TW;
        bool kp = false;
        if (any_surv && (vm[K + li][1] >> lj) & 1u) {
            if (keep_all) {
                kp = true;
            } else {
                // rows outward from the edgel's (dy = 0, -1, 1, -2, ...), and in a
                // row only the survivors: the set bits of its 2K + 1 candidate
                // columns, window mask columns 32 + lj - K .. 32 + lj + K
                const double2 me = posw[(li + K) * WC + (lj + K)];
                const int c0 = TW + lj - K, sh = c0 & 31, wsel = c0 >> 5;
                const uint32_t fmask = (1u << (2 * K + 1)) - 1u;
                for (int t = 0; t <= 2 * K && !kp; ++t) {
                    const int dy = (t & 1) ? -((t + 1) >> 1) : (t >> 1);
                    const int r = K + li + dy;
                    const uint64_t two = ((uint64_t)vm[r][wsel + 1] << 32) | vm[r][wsel];
                    uint32_t bits = (uint32_t)(two >> sh) & fmask;
                    if (dy == 0) bits &= ~(1u << K);          // the edgel itself
                    while (bits) {
                        const int b = __ffs(bits) - 1;
                        bits &= bits - 1;
                        const double2 o = posw[r * WC + lj + b];
                        const double ddx = me.x - o.x, ddy = me.y - o.y;
                        if (ddx * ddx + ddy * ddy <= tp.r2) { kp = true; break; }
                    }
                }
            }
        }
        const unsigned long long b = __ballot(kp);
        if ((lane & 31) == 0) {
            const int row = y0 + 2 * wave + 8 * k + (lane >> 5);
            const uint32_t w = (uint32_t)(b >> lane);
            if (row < H) J.kmask[(long long)row * J.nw + tx] = w;
            // the tile's flag (plain byte stores of one value: no atomics;
            // the scan turns the flags into a list)
            if (w) J.tflag[lt] = 1;
        }
    }
}

// B2: for tiles with kept edgels: the eight planes of the tile + HALO in LDS
// and the cubic sampling / third-order orientation of every kept edgel.
// Three workgroups per CU: E, then Hx and Hy, share the input planes (LDS
// 51.8 KB) and the registers are capped to fit (a couple of spills).
template <int R>
__device__ __forceinline__ void correct_tile(const SpJob &J, const SpTaps &tp, int lt) {
    using G = CGeo<R>;
    constexpr int PR = G::PR, PC = G::PC, IR = G::IR, IC = G::IC;
    __shared__ float lds[G::FLOATS];
    __shared__ uint32_t km[TH];
    __shared__ int s_nlist;
    __shared__ short elist[TH * TW];
    float *In = lds, *X = lds + G::IN, *Y = lds + G::IN + G::XN;

    if (lt >= J.ntx * J.nty) return;
    const int ty = lt / J.ntx, tx = lt - ty * J.ntx;
    const int y0 = ty * TH, x0 = tx * TW;
    const int H = J.H, W = J.W;
    if (threadIdx.x == 0) s_nlist = 0;
    if (threadIdx.x < TH) {
        const int row = y0 + threadIdx.x;
        km[threadIdx.x] = row < H ? J.kmask[(long long)row * J.nw + tx] : 0u;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        const int p = threadIdx.x + k * NT;
        if ((km[p / TW] >> (p % TW)) & 1u) elist[atomicAdd(&s_nlist, 1)] = (short)p;
    }
    __syncthreads();
    if (s_nlist == 0) return;
    // the first round's per-edgel loads (position, strength, rank) are issued
    // now, overlapping the tile load
    constexpr int GQ = 4;                      // lanes per edgel: one cubic stencil row each
    double2 ps_pre = make_double2(x0 + 4.0, y0 + 4.0);
    float2 vv_pre = make_float2(0.f, 0.f);
    int rk_pre = 0;
    if ((int)(threadIdx.x / GQ) < s_nlist) {
        const int p = elist[threadIdx.x / GQ], li = p / TW, lj = p - li * TW;
        const long long o = (long long)(y0 + li) * W + (x0 + lj);
        ps_pre = J.pos[o];
        vv_pre = J.val[o];
        rk_pre = (int)J.wpre[(long long)(y0 + li) * J.nw + tx] + __popc(km[li] & ((1u << lj) - 1u));
    }

    PlaneBox bx;
    {
        int r0 = TH, r1 = -1;
        uint32_t cols = 0u;
#pragma unroll
        for (int i = 0; i < TH; ++i)
            if (km[i]) { r0 = min(r0, i); r1 = i; cols |= km[i]; }
        bx.r0 = r0 & ~1;
        bx.r1 = (r1 + 2 * HALO) | 1;
        bx.c0 = (__ffs(cols) - 1) & ~1;
        bx.c1 = (31 - __clz(cols) + 2 * HALO) | 1;
    }
    load_tile_in<R, 0>(J, In, y0, x0, bx);
    __syncthreads();
    conv_group<R, 0>(In, X, Y, tp, bx);
    load_tile_in<R, 1>(J, In, y0, x0, bx);
    __syncthreads();
    conv_group<R, 1>(In, X, Y, tp, bx);

    // cubic sampling of the ten planes and the orientation: 4 lanes per edgel,
    // lane sa one row of the 4 x 4 stencil (its four pixels), FP64 sums over
    // the quad by DPP; the edgel is written at its raster rank.  Differentiable
    // form: the ten sums and their x and y derivatives (FP32) are kept per
    // edgel for the adjoint, 32 floats at rec[8 rank].
    float4 *out = reinterpret_cast<float4 *>(J.edges);
    const int nlist = s_nlist;
    const int sa = threadIdx.x & (GQ - 1);
#pragma unroll 1
    for (int e0 = 0; e0 < nlist; e0 += NT / GQ) {
        const int e = e0 + (int)(threadIdx.x / GQ);
        const bool act = e < nlist;
        const int p = act ? elist[e] : 0;
        const int li = p / TW, lj = p - li * TW;
        const long long o = (long long)(y0 + li) * W + (x0 + lj);
        double2 ps = ps_pre;
        float2 vv = vv_pre;
        int rank = rk_pre;
        if (e0 > 0) {
            ps = act ? J.pos[o] : make_double2(x0 + 4.0, y0 + 4.0);
            vv = act ? J.val[o] : make_float2(0.f, 0.f);
            rank = act ? (int)J.wpre[(long long)(y0 + li) * J.nw + tx] + __popc(km[li] & ((1u << lj) - 1u)) : 0;
        }
        const int xf = (int)floor(ps.x), yf = (int)floor(ps.y);
        const double fx = ps.x - xf, fy = ps.y - yf;
        const double wya = keys(fy + 1 - sa);
        const int ii = min(max(yf - 1 + sa, 0), H - 1) - (y0 - HALO);
        double v[10];
#pragma unroll
        for (int m = 0; m < 10; ++m) v[m] = 0.0;
        float pv[4][10];
        double wxb[4];
#pragma unroll
        for (int sb = 0; sb < 4; ++sb) {
            wxb[sb] = keys(fx + 1 - sb);
            const int jj = min(max(xf - 1 + sb, 0), W - 1) - (x0 - HALO);
            const int yo = ii * PC + jj, io = (ii + R) * IC + (jj + R);
            pv[sb][0] = Y[0 * PR * PC + yo];   // Px
            pv[sb][1] = Y[1 * PR * PC + yo];   // Py
            pv[sb][2] = Y[2 * PR * PC + yo];   // Pxx
            pv[sb][3] = Y[3 * PR * PC + yo];   // Pyy
            pv[sb][4] = Y[4 * PR * PC + yo];   // Pxy
            pv[sb][5] = In[io];                // Hx
            pv[sb][6] = Y[5 * PR * PC + yo];   // Hxx
            pv[sb][7] = Y[6 * PR * PC + yo];   // Hxy
            pv[sb][8] = Y[7 * PR * PC + yo];   // Hyy
            pv[sb][9] = In[IR * IC + io];      // Hy
            const double w = wya * wxb[sb];
#pragma unroll
            for (int m = 0; m < 10; ++m) v[m] = fma((double)pv[sb][m], w, v[m]);
        }
#pragma unroll
        for (int m = 0; m < 10; ++m) v[m] = toeddpp::group_sum<GQ>(v[m]);
        if (act && sa == 0) {
            double Fx = v[0] * v[6] + v[2] * v[5] + v[1] * v[7] + v[4] * v[9];
            double Fy = v[0] * v[7] + v[4] * v[5] + v[1] * v[8] + v[3] * v[9];
            const double Fm = sqrt(Fx * Fx + Fy * Fy);
            float ori = vv.y;               // atan2(dirx, -diry) from the orientation pass
            if (!(Fm < 1e-12)) {
                // atan2 of F / |F| (:686-690): a positive scale leaves the angle
                // unchanged, so F is scaled exactly by a power of two and the angle
                // taken in FP32, the precision it is stored in
                int ex;
                (void)frexp(fmax(fabs(Fx), fabs(Fy)), &ex);
                ori = atan2f((float)ldexp(Fx, -ex), (float)ldexp(-Fy, -ex));
                J.val[o].y = ori;
            }
            if (rank < J.capacity) {
                // work-grid position / 2^n (exact), :694-695
                out[rank] = make_float4((float)ldexp(ps.x, -J.slog2), (float)ldexp(ps.y, -J.slog2), ori, vv.x);
                if (J.index) J.index[rank] = (int)o;
            }
        }
        if (J.rec) {
            const double dwya = keys_d(fy + 1 - sa);
            float r[32];
#pragma unroll
            for (int m = 0; m < 10; ++m) {
                r[m] = (float)v[m];
                r[10 + m] = 0.f;
                r[20 + m] = 0.f;
            }
            r[30] = __int_as_float((int)o);
            r[31] = 0.f;
#pragma unroll
            for (int sb = 0; sb < 4; ++sb) {
                const float wx = (float)(wya * keys_d(fx + 1 - sb)), wy = (float)(dwya * wxb[sb]);
#pragma unroll
                for (int m = 0; m < 10; ++m) {
                    r[10 + m] = fmaf(pv[sb][m], wx, r[10 + m]);
                    r[20 + m] = fmaf(pv[sb][m], wy, r[20 + m]);
                }
            }
#pragma unroll
            for (int m = 10; m < 30; ++m) r[m] = toeddpp::group_sum<GQ>(r[m]);
            if (act && rank < J.capacity) {
                float4 *dst = J.rec + (long long)rank * 8;
#pragma unroll
                for (int k = 0; k < 8; ++k)
                    if ((k >> 1) == sa) dst[k] = make_float4(r[4 * k], r[4 * k + 1], r[4 * k + 2], r[4 * k + 3]);
            }
        }
    }
}

// A grid of (workgroups, maps) walks each map's list of tiles with kept
// edgels (sp_scan_maps builds it): entries blockIdx.x, blockIdx.x +
// gridDim.x, ..., so every workgroup gets the same number of tiles to within
// one; f(tile) for each, with a barrier between tiles so that a tile's LDS is
// kept until every lane is done with it.
template <class F> __device__ __forceinline__ void walk_kept_tiles(const SpJob &J, F f) {
    const int nl = *J.tlist;
    for (int k = blockIdx.x; k < nl; k += gridDim.x) {
        if (k != (int)blockIdx.x) __syncthreads();
        f(J.tlist[1 + k]);
    }
}

template <int R>
__global__ __launch_bounds__(NT, 3) void sp_correct_kernel(const SpJob *jobs, int njobs, SpTaps tp) {
    const SpJob J = load_job(jobs);
    walk_kept_tiles(J, [&](int t) { correct_tile<R>(J, tp, t); });
}

// ---------------------------------------------------------------------------
// C: scan of kept-word popcounts (raster order), chunk totals, per-map totals
__global__ __launch_bounds__(256) void sp_scan_chunks(const SpJob *jobs, int njobs, int *chunk_sum) {
    __shared__ int wsum[4];
    const SpJob J = load_job(jobs);
    if ((int)blockIdx.x >= J.nchunks) return;
    const int chunk = J.chunk_begin + blockIdx.x;
    const long long nwords = (long long)J.H * J.nw;
    const long long w0 = (long long)blockIdx.x * CHUNKW + threadIdx.x * (CHUNKW / 256);
    int s = 0;
#pragma unroll
    for (int k = 0; k < CHUNKW / 256; ++k)
        if (w0 + k < nwords) s += __popc(J.kmask[w0 + k]);
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) s += __shfl_xor(s, d);
    if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) chunk_sum[chunk] = wsum[0] + wsum[1] + wsum[2] + wsum[3];
}

__global__ __launch_bounds__(1024) void sp_scan_maps(const SpJob *jobs, int *chunk_sum) {
    __shared__ int wsum[16];
    __shared__ int carry, tcarry;
    const SpJob &J = jobs[blockIdx.x];
    if (threadIdx.x == 0) carry = tcarry = 0;
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int base = 0; base < J.nchunks; base += 1024) {
        const int k = base + threadIdx.x;
        const int x = (k < J.nchunks) ? chunk_sum[J.chunk_begin + k] : 0;
        int inc = x;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const int y = __shfl_up(inc, d);
            if (lane >= d) inc += y;
        }
        if (lane == 63) wsum[wave] = inc;
        __syncthreads();
        int wb = carry;
        for (int w = 0; w < wave; ++w) wb += wsum[w];
        if (k < J.nchunks) chunk_sum[J.chunk_begin + k] = wb + inc - x;
        __syncthreads();
        if (threadIdx.x == 1023) carry = wb + inc;
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        if (J.count_out) *J.count_out = carry;
        if (J.gcount) *J.gcount = carry;
        if (carry > J.capacity && J.flags_out) atomicOr(J.flags_out, (int)TOED_FLAG_CAPACITY);
    }
    // the tiles with kept edgels (flags set by the isolation kernel) as a
    // list in raster order, for the kernels that walk it; four flags a thread
    const int ntile = J.ntx * J.nty, nwords = (ntile + 3) / 4;
    for (int base = 0; base < nwords; base += 1024) {
        const int k = base + threadIdx.x;
        uint32_t bits = 0u;
        if (k < nwords) {
            const uint32_t v = reinterpret_cast<const uint32_t *>(J.tflag)[k];
#pragma unroll
            for (int b = 0; b < 4; ++b)
                if (((v >> (8 * b)) & 0xffu) && 4 * k + b < ntile) bits |= 1u << b;
        }
        const int x = __popc(bits);
        int inc = x;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const int y = __shfl_up(inc, d);
            if (lane >= d) inc += y;
        }
        if (lane == 63) wsum[wave] = inc;
        __syncthreads();
        int wb = tcarry;
        for (int w = 0; w < wave; ++w) wb += wsum[w];
        int pos = wb + inc - x;
        while (bits) {
            J.tlist[1 + pos++] = 4 * k + __ffs(bits) - 1;
            bits &= bits - 1u;
        }
        __syncthreads();
        if (threadIdx.x == 1023) tcarry = wb + inc;
        __syncthreads();
    }
    if (threadIdx.x == 0) J.tlist[0] = tcarry;
}

// D: rank of the first kept edgel of every mask word (raster order), from
// the chunk offsets: the correction kernel writes each edgel at its rank
__global__ __launch_bounds__(256) void sp_prefix_kernel(const SpJob *jobs, int njobs, const int *chunk_off) {
    __shared__ int wsum[4];
    const SpJob J = load_job(jobs);
    if ((int)blockIdx.x >= J.nchunks) return;
    const int chunk = J.chunk_begin + blockIdx.x;
    const long long nwords = (long long)J.H * J.nw;
    constexpr int PER = CHUNKW / 256;
    const long long w0 = (long long)blockIdx.x * CHUNKW + threadIdx.x * PER;
    uint32_t wv[PER];
    int s = 0;
#pragma unroll
    for (int k = 0; k < PER; ++k) {
        wv[k] = (w0 + k < nwords) ? J.kmask[w0 + k] : 0u;
        s += __popc(wv[k]);
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int inc = s;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const int y = __shfl_up(inc, d);
        if (lane >= d) inc += y;
    }
    if (lane == 63) wsum[wave] = inc;
    __syncthreads();
    int rank = chunk_off[chunk];
    for (int w = 0; w < wave; ++w) rank += wsum[w];
    rank += inc - s;
#pragma unroll
    for (int k = 0; k < PER; ++k) {
        if (w0 + k < nwords) J.wpre[w0 + k] = (uint32_t)rank;
        rank += __popc(wv[k]);
    }
}

// E: TO_edgemap / TO_orientation by gather, on the IH x IW map grid.  The
// edgel of work pixel (i, j) lies within sqrt 2 of it, so the edgels whose
// rounded map position (x / 2^n, clamped, :708-713) is (Y, X) come from work
// pixels within D = floor(2^n / 2 + sqrt 2) of (2^n Y, 2^n X): the 3 x 3
// neighbourhood when n = 0.
__global__ __launch_bounds__(NT) void sp_maps_kernel(const SpJob *jobs, int njobs) {
    const SpJob J = load_job(jobs);
    const int lt = blockIdx.x;
    const int intx = (J.IW + TW - 1) / TW, inty = (J.IH + TH - 1) / TH;
    if ((!J.to_e && !J.to_o) || lt >= intx * inty) return;
    const int ty = lt / intx, tx = lt - ty * intx;
    const int H = J.H, W = J.W, sc = 1 << J.slog2;
    const int D = (int)floor(0.5 * sc + 1.4142135623730951);
#pragma unroll 1
    for (int k = 0; k < 2; ++k) {
        const int p = threadIdx.x + k * NT;
        const int Y = ty * TH + p / TW, X = tx * TW + (p % TW);
        if (Y >= J.IH || X >= J.IW) continue;
        long long win = -1;
        for (int i = max(sc * Y - D, 0); i <= min(sc * Y + D, H - 1); ++i) {
            for (int j = max(sc * X - D, 0); j <= min(sc * X + D, W - 1); ++j) {
                if (!((J.kmask[(long long)i * J.nw + (j >> 5)] >> (j & 31)) & 1u)) continue;
                const long long o = (long long)i * W + j;
                const double2 ps = J.pos[o];
                const int rx = min(max((int)round(ldexp(ps.x, -J.slog2)), 0), J.IW - 1);
                const int ry = min(max((int)round(ldexp(ps.y, -J.slog2)), 0), J.IH - 1);
                if (rx == X && ry == Y && o > win) win = o;
            }
        }
        float e = 0.f, th = 0.f;
        if (win >= 0) {
            const float2 vv = J.val[win];
            e = vv.x;
            th = vv.y;
        }
        const long long t = (long long)Y * J.tpitch + X;
        if (J.to_e) J.to_e[t] = e;
        if (J.to_o) J.to_o[t] = th;
    }
}

// ---------------------------------------------------------------------------
// Differentiable form: adjoint of the path with every discrete decision
// frozen (edgel set, NMS octant, the half-turn of each pixel's direction,
// sign(-Oxy), cubic stencil base); no reference counterpart.

// NMS octant of a direction, in the reference's branch order (:452-498), the
// offsets (di, dj) of the two pixels of fp and of fm, and the slope family:
// 0 ny/nx, 1 nx/ny, 2 -nx/ny, 3 -ny/nx.
__device__ __forceinline__ int octant_of(double gx, double gy) {
    if (gx >= 0 && gy >= 0) return gx >= gy ? 0 : 1;
    if (gx < 0 && gy >= 0) return fabs(gx) < gy ? 2 : 3;
    if (gx < 0 && gy < 0) return fabs(gx) >= fabs(gy) ? 4 : 5;
    return gx < fabs(gy) ? 6 : 7;
}
// per octant: a (di, dj), b, c, d for fp = Ea (1 - s) + Eb s, fm = Ec (1 - s) + Ed s
__constant__ signed char SP_OFF[8][8] = {
    {0, 1, 1, 1, 0, -1, -1, -1},   {1, 0, 1, 1, -1, 0, -1, -1},  {1, 0, 1, -1, -1, 0, -1, 1},
    {0, -1, 1, -1, 0, 1, -1, 1},   {0, -1, -1, -1, 0, 1, 1, 1},  {-1, 0, -1, -1, 1, 0, 1, 1},
    {-1, 0, -1, 1, 1, 0, 1, -1},   {0, 1, -1, 1, 0, -1, 1, -1}};

template <int R> struct BGeo {
    using G = CGeo<R>;
    static constexpr int PP = G::PR * G::PC;
    static constexpr int XBN = G::IR * G::PC;        // transposed y pass, one X bar plane
    static constexpr int RS = 96;                    // edgels staged per round
    static constexpr int SN = 14 * RS;               // per-edgel staging of one round
    // x-pass tasks of a thread: rows (PR + 2R) x blocks of 4 outputs over PC + 2R columns
    static constexpr int SLOTS = ((G::PR + 2 * R) * ((G::PC + 2 * R + 3) / 4) + NT - 1) / NT;
    static constexpr int NZ = 10;                    // 8 plane adjoints + raw Hx, Hy adjoints
    static constexpr int FLOATS = NZ * PP + (XBN > SN ? XBN : SN);
};

__device__ __forceinline__ float keys_f(float s) {
    const float x = fabsf(s);
    if (x <= 1.f) return (1.5f * x - 2.5f) * x * x + 1.f;
    if (x < 2.f) return ((-0.5f * x + 2.5f) * x - 4.f) * x + 2.f;
    return 0.f;
}

// SB0: one lane per kept edgel, by rank (grid-stride).  From the edgel's
// record (the ten cubic sums and their x, y derivatives, kept by the forward
// pass), its direction and the map values of its NMS stencil, the FP64
// adjoint of the third-order orientation, the position and the NMS parabola:
// dL/dE of the parabola goes to HBM (float atomics), dL/dtheta of the NMS is
// stored at the edgel pixel, and the ten sample adjoints with the cubic
// stencil base and fraction go to brec[4 rank] for the tile kernel.
__global__ __launch_bounds__(NT) void sp_bw_edgel_kernel(const SpJob *jobs, const SpGrad *grads) {
    const SpJob J = load_job(jobs);
    const SpGrad Gd = load_grad(grads);
    const long long n = min((long long)*J.gcount, J.capacity);
    const int W = J.W;
    for (long long rank = (long long)blockIdx.x * NT + threadIdx.x; rank < n; rank += (long long)gridDim.x * NT) {
        const float4 g = Gd.gout[rank];
        const float4 *rp = J.rec + rank * 8;
        float r[32];
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const float4 q = rp[k];
            r[4 * k] = q.x; r[4 * k + 1] = q.y; r[4 * k + 2] = q.z; r[4 * k + 3] = q.w;
        }
        const int o = __float_as_int(r[30]);
        const int gi = o / W, gj = o - gi * W;
        float vb[10];
#pragma unroll
        for (int m = 0; m < 10; ++m) vb[m] = 0.f;
        double thb = 0.0;
        const double2 d = J.dirs[o];
        const float *Em = J.map + (long long)gi * J.pitch + gj;
        auto Ev = [&](int di, int dj) -> double { return (double)Em[(long long)di * J.pitch + dj]; };
        const double md = sqrt(d.x * d.x + d.y * d.y), nx = d.x / md, ny = d.y / md;
        const int oc = octant_of(d.x, d.y);
        const int fam = oc & 3;
        const double slope = fam == 0 ? ny / nx : (fam == 1 ? nx / ny : (fam == 2 ? -nx / ny : -ny / nx));
        const double Ea = Ev(SP_OFF[oc][0], SP_OFF[oc][1]), Eb = Ev(SP_OFF[oc][2], SP_OFF[oc][3]);
        const double Ec = Ev(SP_OFF[oc][4], SP_OFF[oc][5]), Ed = Ev(SP_OFF[oc][6], SP_OFF[oc][7]);
        const double fp = Ea * (1 - slope) + Eb * slope, fm = Ec * (1 - slope) + Ed * slope, f = Ev(0, 0);
        const double s = sqrt(1 + slope * slope);
        const double A = (fm + fp - 2 * f) / (2 * s * s), B = (fp - fm) / (2 * s);
        const double ss = -B / (2 * A), mf = A * ss * ss + B * ss + f;
        const double x = gj + ss * nx, y = gi + ss * ny;
        const float *v = r, *vx = r + 10, *vy = r + 20;
        const double Fx = (double)v[0] * v[6] + (double)v[2] * v[5] + (double)v[1] * v[7] + (double)v[4] * v[9];
        const double Fy = (double)v[0] * v[7] + (double)v[4] * v[5] + (double)v[1] * v[8] + (double)v[3] * v[9];
        const double F2 = Fx * Fx + Fy * Fy;
        if (!(sqrt(F2) < 1e-12)) {
            // orientation = atan2(Fx, -Fy)
            const float Fxb = (float)((double)g.z * (-Fy) / F2), Fyb = (float)((double)g.z * Fx / F2);
            vb[0] = Fxb * v[6] + Fyb * v[7];
            vb[6] = Fxb * v[0];
            vb[2] = Fxb * v[5];
            vb[5] = Fxb * v[2] + Fyb * v[4];
            vb[1] = Fxb * v[7] + Fyb * v[8];
            vb[7] = Fxb * v[1] + Fyb * v[0];
            vb[4] = Fxb * v[9] + Fyb * v[5];
            vb[9] = Fxb * v[4] + Fyb * v[3];
            vb[8] = Fyb * v[1];
            vb[3] = Fyb * v[9];
        } else {
            thb += (double)g.z;   // orientation = atan2(dirx, -diry) = theta + const
        }
        double xb = g.x, yb = g.y;
#pragma unroll
        for (int m = 0; m < 10; ++m) {
            xb += (double)vb[m] * vx[m];
            yb += (double)vb[m] * vy[m];
        }
        // parabola and position: x = j + s* nx, y = i + s* ny, strength |f - B^2 / 4A|
        const double ssb = xb * nx + yb * ny;
        double nxb = xb * ss, nyb = yb * ss;
        const double mfb = (double)g.w * (mf < 0 ? -1.0 : 1.0);
        const double Ab = ssb * (-ss / A) + mfb * ss * ss;
        const double Bb = ssb * (-1.0 / (2 * A)) + mfb * ss;
        const double fb = mfb - Ab / (s * s);
        const double fpb = Ab / (2 * s * s) + Bb / (2 * s), fmb = Ab / (2 * s * s) - Bb / (2 * s);
        const double sbar = Ab * (-2 * A / s) + Bb * (-B / s);
        const double slb = sbar * slope / s + fpb * (Eb - Ea) + fmb * (Ed - Ec);
        // the edgel pixel and the two pixels of fp and of fm
        float *gm = Gd.gmap + (long long)gi * Gd.gpitch + gj;
        atomicAdd(gm, (float)fb);
        atomicAdd(gm + (long long)SP_OFF[oc][0] * Gd.gpitch + SP_OFF[oc][1], (float)(fpb * (1 - slope)));
        atomicAdd(gm + (long long)SP_OFF[oc][2] * Gd.gpitch + SP_OFF[oc][3], (float)(fpb * slope));
        atomicAdd(gm + (long long)SP_OFF[oc][4] * Gd.gpitch + SP_OFF[oc][5], (float)(fmb * (1 - slope)));
        atomicAdd(gm + (long long)SP_OFF[oc][6] * Gd.gpitch + SP_OFF[oc][7], (float)(fmb * slope));
        if (fam == 0) { nxb += -slb * ny / (nx * nx); nyb += slb / nx; }
        else if (fam == 1) { nxb += slb / ny; nyb += -slb * nx / (ny * ny); }
        else if (fam == 2) { nxb += -slb / ny; nyb += slb * nx / (ny * ny); }
        else { nxb += slb * ny / (nx * nx); nyb += -slb / nx; }
        thb += -nxb * ny + nyb * nx;    // n = (cos theta, sin theta)
        J.gth[o] = (float)thb;
        // cubic stencil base (frozen) and fractional position
        const int xfs = (int)floor(x), yfs = (int)floor(y);
        float4 *dst = J.brec + rank * 4;
        dst[0] = make_float4(vb[0], vb[1], vb[2], vb[3]);
        dst[1] = make_float4(vb[4], vb[5], vb[6], vb[7]);
        dst[2] = make_float4(vb[8], vb[9], (float)(x - xfs), (float)(y - yfs));
        dst[3] = make_float4(__int_as_float(xfs), __int_as_float(yfs), 0.f, 0.f);
    }
}

// SB1: per tile with kept edgels.  (a) one lane per edgel: its record from
// the edgel kernel is staged in LDS, and the Z rows and columns its cubic
// stencil reaches widen the tile's box.  (b) 16 lanes per edgel, lane (a, b)
// one pixel of the 4 x 4 stencil: the eight convolved-plane adjoints and the
// raw Hx, Hy sample adjoints are scattered into Z (LDS atomics).  Then the
// transposed circular convolutions pull Z back to dL/dE, dL/dHx, dL/dHy over
// the tile + HALO + R (the raw adjoints join dL/dHx, dL/dHy there); through
// Hx = E dirx, Hy = E diry they become dL/dE and dL/dtheta at each pixel,
// added to HBM with float atomics (periodic wrap).
template <int R>
__device__ __forceinline__ void bw_tile(const SpJob &J, const SpGrad &Gd, const SpTaps &tp, int lt) {
    using G = CGeo<R>;
    constexpr int PR = G::PR, PC = G::PC, IR = G::IR, IC = G::IC, PP = PR * PC;
    __shared__ float lds[BGeo<R>::FLOATS];
    __shared__ uint32_t km[TH];
    __shared__ int s_any, s_nlist;
    __shared__ int s_box[4];           // rows, cols of Z that the edgels' stencils reach: min, max
    __shared__ short elist[TH * TW];
    // Z (the 8 plane adjoints, then the raw Hx, Hy sample adjoints) | Xb
    // (transposed y pass), which holds the per-edgel staging St during the
    // edgel phase
    constexpr int NZ = BGeo<R>::NZ;
    float *Z = lds, *Xb = lds + NZ * PP, *St = lds + NZ * PP;

    if (lt >= J.ntx * J.nty) return;
    const int ty = lt / J.ntx, tx = lt - ty * J.ntx;
    const int y0 = ty * TH, x0 = tx * TW;
    const int H = J.H, W = J.W;
    if (threadIdx.x == 0) { s_any = 0; s_nlist = 0; s_box[0] = PR; s_box[1] = -1; s_box[2] = PC; s_box[3] = -1; }
    __syncthreads();
    if (threadIdx.x < TH) {
        const int row = y0 + threadIdx.x;
        const uint32_t w = row < H ? J.kmask[(long long)row * J.nw + tx] : 0u;
        km[threadIdx.x] = w;
        if (w) s_any = 1;
    }
    __syncthreads();
    if (!s_any) return;
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        const int p = threadIdx.x + k * NT;
        if ((km[p / TW] >> (p % TW)) & 1u) elist[atomicAdd(&s_nlist, 1)] = (short)p;
    }
    for (int idx = threadIdx.x; idx < NZ * PP; idx += NT) Z[idx] = 0.f;
    __syncthreads();

    // v index -> adjoint plane: Px 0, Py 1, Pxx 2, Pyy 3, Pxy 4, Hxx 5, Hxy 6, Hyy 7,
    // raw Hx 8, raw Hy 9
    constexpr int VPL[10] = {0, 1, 2, 3, 4, 8, 5, 6, 7, 9};
    const int nlist = s_nlist;
    const int sub = threadIdx.x & 15, sa = sub >> 2, sb = sub & 3;
    constexpr int RS = BGeo<R>::RS;
#pragma unroll 1
    for (int e0 = 0; e0 < nlist; e0 += RS) {
        // (a) one lane per edgel: its staged record, or nothing past the capacity
        if (threadIdx.x < RS) {
            const int e = e0 + threadIdx.x;
            float4 q0 = make_float4(0.f, 0.f, 0.f, 0.f), q1 = q0, q2 = q0;
            int xfs = x0, yfs = y0;
            if (e < nlist) {
                const int p = elist[e], li = p / TW, lj = p - li * TW;
                const int gi = y0 + li, gj = x0 + lj;
                const int rank = (int)J.wpre[(long long)gi * J.nw + tx] + __popc(km[li] & ((1u << lj) - 1u));
                if (rank < J.capacity) {
                    const float4 *bp = J.brec + (long long)rank * 4;
                    q0 = bp[0];
                    q1 = bp[1];
                    q2 = bp[2];
                    const float4 q3 = bp[3];
                    xfs = __float_as_int(q3.x);
                    yfs = __float_as_int(q3.y);
                    // the Z rows and columns the stencil reaches (borders clamped)
                    atomicMin(&s_box[0], max(yfs - 1, 0) - (y0 - HALO));
                    atomicMax(&s_box[1], min(yfs + 2, H - 1) - (y0 - HALO));
                    atomicMin(&s_box[2], max(xfs - 1, 0) - (x0 - HALO));
                    atomicMax(&s_box[3], min(xfs + 2, W - 1) - (x0 - HALO));
                }   // else not in the output: no gradient (its dL/dtheta stays cleared)
            }
            const float vb[10] = {q0.x, q0.y, q0.z, q0.w, q1.x, q1.y, q1.z, q1.w, q2.x, q2.y};
#pragma unroll
            for (int m = 0; m < 10; ++m) St[m * RS + threadIdx.x] = vb[m];
            St[10 * RS + threadIdx.x] = q2.z;
            St[11 * RS + threadIdx.x] = q2.w;
            St[12 * RS + threadIdx.x] = __int_as_float(xfs);
            St[13 * RS + threadIdx.x] = __int_as_float(yfs);
        }
        __syncthreads();
        // (b) 16 lanes per edgel: lane (sa, sb) scatters its stencil pixel
        const int nr = min(RS, nlist - e0);
#pragma unroll 1
        for (int q = threadIdx.x >> 4; q < nr; q += NT / 16) {
            const float fx = St[10 * RS + q], fy = St[11 * RS + q];
            const int xf = __float_as_int(St[12 * RS + q]), yf = __float_as_int(St[13 * RS + q]);
            const float w = keys_f(fy + 1 - sa) * keys_f(fx + 1 - sb);
            const int gy = min(max(yf - 1 + sa, 0), H - 1), gx = min(max(xf - 1 + sb, 0), W - 1);
            const int yo = (gy - (y0 - HALO)) * PC + (gx - (x0 - HALO));
#pragma unroll
            for (int m = 0; m < 10; ++m) {
                const float c = St[m * RS + q] * w;
                if (c != 0.f) atomicAdd(&Z[VPL[m] * PP + yo], c);
            }
        }
        __syncthreads();
    }

    // transposed y pass (Z -> X bar), then x pass (X bar -> dL/dE, dL/dHx,
    // dL/dHy, then dL/dE and dL/dtheta to HBM), per group and one X bar plane at a time: the x
    // pass adds each plane's share into per-thread registers (the tasks of a
    // thread are the same for the three planes), so only one X bar plane is
    // held in LDS.  Output u of a 4-output block at o0 sums window[u + m] *
    // k[m], m = 0..2R, with window[t] = source(o0 - 2R + t) (zero outside the
    // source): the transposed filters.
    // Only Z rows rlo .. rmax and columns clo .. chi can be nonzero: X bar rows
    // rlo .. rlo + nrow - 1 (= rmax + 2R) and outputs up to column chi + 2R.
    constexpr int NB = 4, NWIN = 2 * R + NB;
    constexpr int SLOTS = BGeo<R>::SLOTS;
    const int rlo = s_box[0], clo = s_box[2], chi = s_box[3];
    if (s_box[1] < rlo) return;                       // no stencil reached Z
    const int nrow = s_box[1] + 2 * R - rlo + 1, ncol = chi - clo + 1;
    const int nyb = (nrow + NB - 1) / NB, nxb = (ncol + 2 * R + NB - 1) / NB;
    const int iy0 = y0 - HALO - R, ix0 = x0 - HALO - R;
    const bool inner = iy0 >= 0 && iy0 + IR <= H && ix0 >= 0 && ix0 + IC <= W;
    // (adjoint plane, X bar plane, filter family) of each group
    constexpr int NSRC = 5;
    constexpr int src[2][NSRC] = {{0, 4, 1, 3, 2}, {5, 7, 6, -1, -1}};
    constexpr int dst[2][NSRC] = {{0, 0, 1, 1, 2}, {0, 1, 2, -1, -1}};
    constexpr int fam[2][NSRC] = {{0, 1, 1, 2, 0}, {0, 1, 0, -1, -1}};
    float ge[SLOTS][NB];      // group 0's dL/dE, added with group 1's terms at the end
#pragma unroll
    for (int grp = 0; grp < 2; ++grp) {
        float ex[SLOTS][2][NB];
#pragma unroll
        for (int sl = 0; sl < SLOTS; ++sl)
#pragma unroll
            for (int u = 0; u < NB; ++u) ex[sl][0][u] = ex[sl][1][u] = 0.f;
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            for (int task = threadIdx.x; task < nyb * ncol; task += NT) {
                const int blk = task / ncol, c = clo + (task - blk * ncol), r0 = rlo + blk * NB;
                float acc[NB];
#pragma unroll
                for (int u = 0; u < NB; ++u) acc[u] = 0.f;
#pragma unroll
                for (int s = 0; s < NSRC; ++s) {
                    if (src[grp][s] < 0 || dst[grp][s] != q) continue;
                    const float *zp = Z + src[grp][s] * PP + c;
                    float win[NWIN];
#pragma unroll
                    for (int t = 0; t < NWIN; ++t) {
                        const int r = r0 - 2 * R + t;
                        win[t] = (r >= 0 && r < PR) ? zp[r * PC] : 0.f;
                    }
                    const int f = fam[grp][s];
#pragma unroll
                    for (int m = 0; m <= 2 * R; ++m) {
                        const float kk = tp.k[f][m];
#pragma unroll
                        for (int u = 0; u < NB; ++u) acc[u] = fmaf(win[u + m], kk, acc[u]);
                    }
                }
#pragma unroll
                for (int u = 0; u < NB; ++u)
                    if (r0 + u < rlo + nrow) Xb[(r0 + u) * PC + c] = acc[u];
            }
            __syncthreads();
            // group 0: E bar = X0 * G' + X1 * G + X2 * G''; group 1: Hx bar = X0 * G',
            // Hy bar = X1 * G + X2 * G'
            const int f = grp == 0 ? (q == 0 ? 1 : (q == 1 ? 0 : 2)) : (q == 0 ? 1 : (q == 1 ? 0 : 1));
            const int to1 = (grp == 1 && q > 0) ? 1 : 0;
#pragma unroll
            for (int sl = 0; sl < SLOTS; ++sl) {
                const int task = threadIdx.x + sl * NT;
                if (task >= nrow * nxb) continue;
                const int rr = task / nxb, blk = task - rr * nxb, r = rlo + rr, c0 = clo + blk * NB;
                const float *xp = Xb + r * PC;
                float win[NWIN];
#pragma unroll
                for (int t = 0; t < NWIN; ++t) {
                    const int c = c0 - 2 * R + t;
                    win[t] = (c >= clo && c <= chi) ? xp[c] : 0.f;
                }
#pragma unroll
                for (int m = 0; m <= 2 * R; ++m) {
                    const float kk = tp.k[f][m];
#pragma unroll
                    for (int u = 0; u < NB; ++u) ex[sl][to1][u] = fmaf(win[u + m], kk, ex[sl][to1][u]);
                }
            }
            __syncthreads();      // the next plane's y pass overwrites X bar
        }
#pragma unroll
        for (int sl = 0; sl < SLOTS; ++sl) {
            const int task = threadIdx.x + sl * NT;
            if (task >= nrow * nxb) continue;
            const int rr = task / nxb, blk = task - rr * nxb, r = rlo + rr, c0 = clo + blk * NB;
            const int gy = inner ? iy0 + r : wrap_circ(iy0 + r, H);
#pragma unroll
            for (int u = 0; u < NB; ++u) {
                const int cm = c0 + u;
                if (cm > chi + 2 * R) continue;
                const int gx = inner ? ix0 + cm : wrap_circ(ix0 + cm, W);
                float e0 = ex[sl][0][u], e1 = ex[sl][1][u];
                if (grp == 1) {
                    // the raw Hx, Hy sample adjoints of the cubic stencils (plane
                    // coordinates = output coordinates - R)
                    const int zr = r - R, zc = cm - R;
                    if (zr >= 0 && zr < PR && zc >= 0 && zc < PC) {
                        e0 += Z[8 * PP + zr * PC + zc];
                        e1 += Z[9 * PP + zr * PC + zc];
                    }
                }
                if (grp == 0) {
                    ge[sl][u] = e0;
                } else if (ge[sl][u] != 0.f || e0 != 0.f || e1 != 0.f) {
                    // Hx = E dirx, Hy = E diry at this pixel: dL/dHx = e0 and
                    // dL/dHy = e1 go to dL/dE (dirx e0 + diry e1) and to
                    // dL/dtheta (E (dirx e1 - diry e0)) here, so only two
                    // planes take atomics
                    const float4 q = J.tq[(long long)gy * W + gx];
                    const float ev = J.map[(long long)gy * J.pitch + gx];
                    const float de = ge[sl][u] + q.x * e0 + q.y * e1, dt = ev * (q.x * e1 - q.y * e0);
                    if (de != 0.f) atomicAdd(&Gd.gmap[(long long)gy * Gd.gpitch + gx], de);
                    if (dt != 0.f) atomicAdd(&J.gth[(long long)gy * W + gx], dt);
                }
            }
        }
    }
}

// Walks each map's set of tiles with kept edgels, as the correction does.
template <int R>
__global__ __launch_bounds__(NT) void sp_bw_tile_kernel(const SpJob *jobs, const SpGrad *grads, SpTaps tp) {
    const SpJob J = load_job(jobs);
    const SpGrad Gd = load_grad(grads);
    walk_kept_tiles(J, [&](int t) { bw_tile<R>(J, Gd, tp, t); });
}

// SB2: dense, per tile.  From dL/dtheta at every pixel (the edgel's own NMS
// term and, added by SB1, that of Hx = E cos(theta), Hy = E sin(theta)),
// dL/dOxx and dL/dOyy through theta = atan(Oyy sgn / (Oxx + 1e-5)), then the
// transposed central differences (one-sided at the border) and the transposed
// triangle smoothing (half-sample-symmetric border) into dL/dE.
constexpr int DR = TH + 12, DC = TW + 12;    // tile +- 6

__device__ __forceinline__ float d_transpose(const float *b, int stride, int j, int n, int jo) {
    // (D^T b)(j), D the central difference with one-sided ends; b(k) = b[(k - jo) * stride]
    float acc = 0.f;
    if (j >= 2 && j <= n - 3) {
        // interior: only the central terms reach j (same operations, same order)
        acc += 0.5f * b[(j - 1 - jo) * stride];
        acc -= 0.5f * b[(j + 1 - jo) * stride];
        return acc;
    }
    if (j - 1 >= 1 && j - 1 <= n - 2) acc += 0.5f * b[(j - 1 - jo) * stride];
    if (j + 1 >= 1 && j + 1 <= n - 2) acc -= 0.5f * b[(j + 1 - jo) * stride];
    if (j == 1) acc += b[(0 - jo) * stride];
    if (j == 0) acc -= b[(0 - jo) * stride];
    if (j == n - 1) acc += b[(n - 1 - jo) * stride];
    if (j == n - 2) acc -= b[(n - 1 - jo) * stride];
    return acc;
}

__device__ __forceinline__ float tri_transpose(const float *b, int stride, int m, int n, int jo) {
    // (T^T b)(m) for T the 9-tap triangle with half-sample-symmetric ends (n >= 5)
    float acc = 0.f;
    if (m >= 4 && m <= n - 5) {
        // interior: no folded term reaches m (same operations, same order)
#pragma unroll
        for (int q = -4; q <= 4; ++q) acc += (float)(5 - (q < 0 ? -q : q)) / 25.f * b[(m + q - jo) * stride];
        return acc;
    }
#pragma unroll
    for (int q = -4; q <= 4; ++q) {
        const float t = (float)(5 - (q < 0 ? -q : q)) / 25.f;
        const int i = m + q;
        if (i >= 0 && i < n) acc += t * b[(i - jo) * stride];
        const int il = q - 1 - m;             // i - q = -1 - m folds to m
        if (il >= 0 && il < n) acc += t * b[(il - jo) * stride];
        const int ir = 2 * n - 1 - m + q;     // i - q = 2n - 1 - m folds to m
        if (ir >= 0 && ir < n) acc += t * b[(ir - jo) * stride];
    }
    return acc;
}

__global__ __launch_bounds__(NT) void sp_bw_dense_kernel(const SpJob *jobs, const SpGrad *grads, int R) {
    __shared__ float Ox[DR][DC], Oy[DR][DC], T1[DR][DC], T2[DR][DC];
    const SpJob J = load_job(jobs);
    const int lt = blockIdx.x;
    if (lt >= J.ntx * J.nty) return;
    const SpGrad Gd = load_grad(grads);
    const int ty = lt / J.ntx, tx = lt - ty * J.ntx;
    const int y0 = ty * TH, x0 = tx * TW;
    const int H = J.H, W = J.W;
    const int ry0 = y0 - 6, rx0 = x0 - 6;
    {
        // Every input of this tile is zero unless a kept edgel lies within
        // 6 + HALO + R pixels of it: dL/dtheta reaches HALO + R from an
        // edgel's pixel (cubic stencil, transposed filters), dL/dtheta of the
        // NMS sits at the edgel, and the transposed differences and smoothing
        // read 6 more.  Such interior tiles leave dL/dE unchanged; tiles near
        // the border always run (the filters wrap around the map).
        const int reach = 6 + HALO + R;
        if (y0 - reach >= 0 && y0 + TH + reach <= H && x0 - reach >= 0 && x0 + TW + reach <= W) {
            const int r0 = y0 - reach, nr = TH + 2 * reach;
            const int c0 = x0 - reach, c1 = x0 + TW - 1 + reach;
            const int w0 = c0 >> 5, nwd = (c1 >> 5) - w0 + 1;
            int hit = 0;
            for (int k = threadIdx.x; k < nr * nwd; k += NT) {
                const int r = r0 + k / nwd, wi = w0 + k % nwd;
                const int lo = max(c0 - 32 * wi, 0), hi = min(c1 - 32 * wi, 31);
                const uint32_t keep = (hi == 31 ? 0xffffffffu : ((1u << (hi + 1)) - 1u)) & ~((1u << lo) - 1u);
                hit |= (J.kmask[(long long)r * J.nw + wi] & keep) != 0u;
            }
            if (!__syncthreads_or(hit)) return;
        }
    }
    {
        // all of a thread's region loads are issued before any is used
        constexpr int NU = (DR * DC + NT - 1) / NT;
        float4 qv[NU];
        float gtv[NU];
        bool in[NU];
#pragma unroll
        for (int u = 0; u < NU; ++u) {
            const int idx = threadIdx.x + u * NT;
            const int r = idx / DC, c = idx - r * DC;
            const int gy = ry0 + r, gx = rx0 + c;
            in[u] = idx < DR * DC && gy >= 0 && gy < H && gx >= 0 && gx < W;
            qv[u] = make_float4(0.f, 0.f, 0.f, 0.f);
            gtv[u] = 0.f;
            if (in[u]) {
                const long long o = (long long)gy * W + gx;
                qv[u] = J.tq[o];
                gtv[u] = J.gth[o];
            }
        }
#pragma unroll
        for (int u = 0; u < NU; ++u) {
            const int idx = threadIdx.x + u * NT;
            if (idx >= DR * DC) continue;
            const int r = idx / DC, c = idx - r * DC;
            float ox = 0.f, oy = 0.f;
            if (in[u]) {
                const float4 q = qv[u];
                ox = gtv[u] * q.z;
                oy = gtv[u] * q.w;
            }
            Ox[r][c] = ox;
            Oy[r][c] = oy;
        }
    }
    __syncthreads();
    // one transposed difference: x on Ox (cols 1..DC-2), y on Oy (rows 1..DR-2)
    for (int idx = threadIdx.x; idx < DR * DC; idx += NT) {
        const int r = idx / DC, c = idx - r * DC;
        const int gy = ry0 + r, gx = rx0 + c;
        float a = 0.f, b = 0.f;
        if (gy >= 0 && gy < H && gx >= 0 && gx < W) {
            if (c >= 1 && c <= DC - 2) a = d_transpose(&Ox[r][0], 1, gx, W, rx0);
            if (r >= 1 && r <= DR - 2) b = d_transpose(&Oy[0][c], DC, gy, H, ry0);
        }
        T1[r][c] = a;
        T2[r][c] = b;
    }
    __syncthreads();
    // second transposed difference, summed: S bar on rows/cols 2..DR-3 / DC-3
    for (int idx = threadIdx.x; idx < DR * DC; idx += NT) {
        const int r = idx / DC, c = idx - r * DC;
        const int gy = ry0 + r, gx = rx0 + c;
        float sb = 0.f;
        if (gy >= 0 && gy < H && gx >= 0 && gx < W && r >= 2 && r <= DR - 3 && c >= 2 && c <= DC - 3)
            sb = d_transpose(&T1[r][0], 1, gx, W, rx0) + d_transpose(&T2[0][c], DC, gy, H, ry0);
        Ox[r][c] = sb;
    }
    __syncthreads();
    // transposed column pass (rows of the tile, cols 2..DC-3)
    for (int idx = threadIdx.x; idx < TH * DC; idx += NT) {
        const int rr = idx / DC, c = idx - rr * DC, r = rr + 6;
        const int gy = ry0 + r, gx = rx0 + c;
        float v = 0.f;
        if (gy < H && gx >= 0 && gx < W && c >= 2 && c <= DC - 3) v = tri_transpose(&Ox[0][c], DC, gy, H, ry0);
        Oy[r][c] = v;
    }
    __syncthreads();
    // transposed row pass at the tile pixels, plus dL/dE of Hx = E dirx, Hy = E diry
#pragma unroll
    for (int k = 0; k < TH * TW / NT; ++k) {
        const int p = threadIdx.x + k * NT;
        const int li = p / TW, lj = p - li * TW;
        const int gi = y0 + li, gj = x0 + lj;
        if (gi >= H || gj >= W) continue;
        const float ge = tri_transpose(&Oy[li + 6][0], 1, gj, W, rx0);
        float *dst = &Gd.gmap[(long long)gi * Gd.gpitch + gj];
        *dst += ge;
    }
}

// zero dL/dtheta of every map (a kernel: the sequence stays capturable)
__global__ void sp_bw_clear_kernel(const SpJob *jobs) {
    const SpJob J = load_job(jobs);
    const long long n = (long long)J.H * J.W;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        J.gth[i] = 0.f;
    }
}

// zero per-map counters, flags and kept-tile bits (a kernel, so the sequence
// can be captured); grid (bit words / 256, maps)
__global__ void sp_reset_kernel(const SpJob *jobs) {
    const SpJob J = load_job(jobs);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        *J.nsurv = 0;
        if (J.flags_out) *J.flags_out = 0;
        if (J.count_out) *J.count_out = 0;
    }
    const int nwords = (J.ntx * J.nty + 3) / 4;      // the flag bytes, four at a time
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nwords; i += gridDim.x * blockDim.x)
        reinterpret_cast<uint32_t *>(J.tflag)[i] = 0u;
}

// R: interp_n > 0 only: the caller's IH x IW map (and given orientation) to
// the H x W work grid by bilinear interpolation with the corners aligned
// (:69-90, set_edgemap / set_orientation :226-240): work pixel (i, j) samples
// (i (IH - 1) / (H - 1), j (IW - 1) / (W - 1)), the lower / right neighbour
// clamped to the last row / column.  FP64 in the reference's operation order
// (contraction off, correctly rounded division), stored as float.
__global__ __launch_bounds__(NT) void sp_resize_kernel(const SpJob *jobs, int njobs) {
#pragma clang fp contract(off)
    const SpJob J = load_job(jobs);
    const int lt = blockIdx.x;
    if (J.slog2 == 0 || lt >= J.ntx * J.nty) return;
    const int ty = lt / J.ntx, tx = lt - ty * J.ntx;
    float *dmap = const_cast<float *>(J.map);
    float *dori = const_cast<float *>(J.orient);
#pragma unroll 1
    for (int k = 0; k < 2; ++k) {
        const int p = threadIdx.x + k * NT;
        const int i = ty * TH + p / TW, j = tx * TW + (p % TW);
        if (i >= J.H || j >= J.W) continue;
        const double y = (double)i * (J.IH - 1) / (double)(J.H - 1);
        const double x = (double)j * (J.IW - 1) / (double)(J.W - 1);
        const int ya = (int)floor(y), yb = min(ya + 1, J.IH - 1);
        const int xa = (int)floor(x), xb = min(xa + 1, J.IW - 1);
        const double fy = y - ya, fx = x - xa;
        auto bil = [&](const float *src, int pitch) -> float {
            const float *ra = src + (long long)ya * pitch, *rb = src + (long long)yb * pitch;
            const double v = (1 - fy) * ((1 - fx) * (double)ra[xa] + fx * (double)ra[xb]) +
                             fy * ((1 - fx) * (double)rb[xa] + fx * (double)rb[xb]);
            return (float)v;
        };
        dmap[(long long)i * J.W + j] = bil(J.src_map, J.spitch);
        if (J.src_orient) dori[(long long)i * J.W + j] = bil(J.src_orient, J.sopitch);
    }
}

}  // namespace toedsp

using namespace toedsp;

struct toed_subpix {
    toed_ctx *ctx = nullptr;
    int device = 0, n = 0, R = 0;
    SpTaps taps{};
    std::vector<SpJob> jobs;
    SpJob *d_jobs = nullptr;
    char *d_work = nullptr;
    int *d_chunk = nullptr;
    int *d_nsurv = nullptr;
    long long ntiles = 0, nchunks = 0;
    int max_tiles = 0, max_chunks = 0;
    int slog2 = 0;              // interp_n: the maps run on their 2^n work grids
    bool timing = false;
    hipEvent_t ev[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    // differentiable form
    bool grad = false, grad_ran = false;
    bool dirty = false;         // maps re-pointed on the host since the last upload
    char *d_gwork = nullptr;
    std::vector<size_t> goff;
    int eblocks = 0;            // edgel-kernel workgroups per map
    SpGrad *d_grads = nullptr;
};

#define SPCHK(x)                                                                                         \
    do {                                                                                                 \
        hipError_t e_ = (x);                                                                             \
        if (e_ != hipSuccess) {                                                                          \
            std::fprintf(stderr, "toed: %s failed: %s (%s:%d)\n", #x, hipGetErrorString(e_), __FILE__, __LINE__); \
            return TOED_ERR_HIP;                                                                         \
        }                                                                                                \
    } while (0)

// workgroups per map of the kernels that walk the kept-tile lists: about
// SP_WALK_WGS in all, never more than a map's tiles
constexpr int SP_WALK_WGS = 4096;
static int walk_grid(const toed_subpix *p) {
    return std::max(1, std::min(p->max_tiles, (SP_WALK_WGS + p->n - 1) / p->n));
}

template <int R> static void launch_bw_tile(hipStream_t s, const toed_subpix *p) {
    hipLaunchKernelGGL(sp_bw_tile_kernel<R>, dim3(walk_grid(p), p->n), dim3(NT), 0, s, p->d_jobs, p->d_grads,
                       p->taps);
}

static int launch_bw_tile_r(hipStream_t s, const toed_subpix *p) {
    switch (p->R) {
#define SP_CASE(r) case r: launch_bw_tile<r>(s, p); break;
        SP_CASE(1) SP_CASE(2) SP_CASE(3) SP_CASE(4) SP_CASE(5) SP_CASE(6)
        SP_CASE(7) SP_CASE(8) SP_CASE(9) SP_CASE(10) SP_CASE(11) SP_CASE(12)
#undef SP_CASE
        default: return TOED_ERR_SIGMA;
    }
    return TOED_OK;
}

template <int R> static void launch_correct(hipStream_t s, const toed_subpix *p) {
    hipLaunchKernelGGL(sp_correct_kernel<R>, dim3(walk_grid(p), p->n), dim3(NT), 0, s, p->d_jobs, p->n, p->taps);
}

static int launch_correct_r(hipStream_t s, const toed_subpix *p) {
    switch (p->R) {
#define SP_CASE(r) case r: launch_correct<r>(s, p); break;
        SP_CASE(1) SP_CASE(2) SP_CASE(3) SP_CASE(4) SP_CASE(5) SP_CASE(6)
        SP_CASE(7) SP_CASE(8) SP_CASE(9) SP_CASE(10) SP_CASE(11) SP_CASE(12)
#undef SP_CASE
        default: return TOED_ERR_SIGMA;
    }
    return TOED_OK;
}

static size_t align_up(size_t v) { return (v + 255) & ~(size_t)255; }

extern "C" {

int toed_subpix_radius(double sigma, int *R) {
    if (!R) return TOED_ERR_ARG;
    if (!(sigma > 0) || !std::isfinite(sigma)) return TOED_ERR_SIGMA;
    const int r = (int)std::ceil(sigma * 4);
    if (r < 1 || r > MAXR) return TOED_ERR_SIGMA;
    *R = r;
    return TOED_OK;
}

int toed_subpix_plan(toed_ctx *ctx, int n, const toed_subpix_map_t *maps, double sigma, double nbr_radius,
                     toed_subpix **out) {
    return toed_subpix_plan_n(ctx, n, maps, sigma, nbr_radius, 0, out);
}

int toed_subpix_plan_n(toed_ctx *ctx, int n, const toed_subpix_map_t *maps, double sigma, double nbr_radius,
                       int interp_n, toed_subpix **out) {
    if (!ctx || !out || n <= 0 || n > 65535 || !maps || interp_n < 0 || interp_n > TOED_SUBPIX_MAX_INTERP)
        return TOED_ERR_ARG;
    *out = nullptr;
    int R = 0;
    const int st = toed_subpix_radius(sigma, &R);
    if (st) return st;
    if (!(nbr_radius == nbr_radius)) return TOED_ERR_ARG;
    const int sc = 1 << interp_n;
    // the isolation radius on the work grid: positions there are 2^n times the
    // list's, so distances compare exactly the same way (powers of two)
    const double rw = nbr_radius * sc;
    const int K = nbr_radius > 0 ? (int)std::floor(rw + 2.8284271247461903 + 1e-9) : 0;
    if (K > MAXK_WIDE) return TOED_ERR_ARG;
    auto work = [sc](int d) { return (long long)(d - 1) * sc + 1; };
    for (int m = 0; m < n; ++m) {
        const toed_subpix_map_t &M = maps[m];
        if (!M.d_map || M.H < 2 || M.W < 2 || M.pitch < M.W || !M.d_edges || M.capacity < 0) return TOED_ERR_ARG;
        if (M.d_orient && M.opitch < M.W) return TOED_ERR_ARG;
        if ((M.d_to_edgemap || M.d_to_orientation) && M.tpitch < M.W) return TOED_ERR_ARG;
        if (work(M.H) * work(M.W) >= (1LL << 31)) return TOED_ERR_ARG;
    }
    toed_subpix *p = new toed_subpix();
    p->ctx = ctx;
    p->device = toed_get_device(ctx);
    p->n = n;
    p->R = R;
    p->slog2 = interp_n;
    // taps (:194-214), FP64 closed forms rounded once to FP32
    {
        const double s2 = sigma * sigma, s3 = s2 * sigma, s5 = s2 * s3, nrm = std::sqrt(2.0 * M_PI);
        for (int k = -R; k <= R; ++k) {
            const double e = std::exp(-((double)k * k) / (2 * s2));
            p->taps.k[0][k + R] = (float)(e / (nrm * sigma));
            p->taps.k[1][k + R] = (float)(-k * e / (nrm * s3));
            p->taps.k[2][k + R] = (float)(((double)k * k - s2) * e / (nrm * s5));
        }
        p->taps.R = R;
        p->taps.K = K;
        p->taps.keep_all = nbr_radius > 0 ? 0 : 1;
        p->taps.r2 = rw * rw;
    }
    // work planes per map (and, with interp_n > 0, the resized map and orientation)
    size_t total = 0;
    std::vector<size_t> off(n);
    for (int m = 0; m < n; ++m) {
        const size_t wh = (size_t)work(maps[m].H), ww = (size_t)work(maps[m].W);
        const size_t px = wh * ww, nw = (ww + 31) / 32;
        off[m] = total;
        const size_t ntile = nw * ((wh + TH - 1) / TH);
        total += align_up(px * 4) * 2 + align_up(px * 16) + align_up(px * 8) + align_up(wh * nw * 4) * 3 +
                 align_up(ntile) + align_up((ntile + 1) * 4);
        if (interp_n) total += align_up(px * 4) * (maps[m].d_orient ? 2 : 1);
    }
    if (hipSetDevice(p->device) != hipSuccess) { delete p; return TOED_ERR_HIP; }
    if (hipMalloc(&p->d_work, total) != hipSuccess) { delete p; return TOED_ERR_ALLOC; }
    if (hipMalloc(&p->d_nsurv, 256 * (size_t)n) != hipSuccess) { toed_subpix_destroy(p); return TOED_ERR_ALLOC; }
    long long tiles = 0, chunks = 0;
    p->jobs.resize(n);
    for (int m = 0; m < n; ++m) {
        const toed_subpix_map_t &M = maps[m];
        SpJob &J = p->jobs[m];
        std::memset(&J, 0, sizeof(J));
        const int wh = (int)work(M.H), ww = (int)work(M.W);
        const size_t px = (size_t)wh * ww;
        char *b = p->d_work + off[m];
        J.hx = (float *)b; b += align_up(px * 4);
        J.hy = (float *)b; b += align_up(px * 4);
        J.pos = (double2 *)b; b += align_up(px * 16);
        J.val = (float2 *)b; b += align_up(px * 8);
        J.nw = (ww + 31) / 32;
        J.vmask = (uint32_t *)b; b += align_up((size_t)wh * J.nw * 4);
        J.kmask = (uint32_t *)b; b += align_up((size_t)wh * J.nw * 4);
        J.wpre = (uint32_t *)b; b += align_up((size_t)wh * J.nw * 4);
        J.tflag = (uint8_t *)b; b += align_up((size_t)J.nw * ((wh + TH - 1) / TH));
        J.tlist = (int *)b; b += align_up(((size_t)J.nw * ((wh + TH - 1) / TH) + 1) * 4);
        J.nsurv = p->d_nsurv + 64 * m;
        J.src_map = M.d_map;
        J.src_orient = M.d_orient;
        J.IH = M.H;
        J.IW = M.W;
        J.spitch = M.pitch;
        J.sopitch = M.d_orient ? M.opitch : M.W;
        J.slog2 = interp_n;
        J.margin = 3 * sc;
        if (interp_n) {
            J.map = (const float *)b; b += align_up(px * 4);
            J.orient = M.d_orient ? (const float *)b : nullptr;
            J.pitch = J.opitch = ww;
        } else {
            J.map = M.d_map;
            J.orient = M.d_orient;
            J.pitch = M.pitch;
            J.opitch = M.d_orient ? M.opitch : M.W;
        }
        J.edges = M.d_edges;
        J.index = M.d_index;
        J.to_e = M.d_to_edgemap;
        J.to_o = M.d_to_orientation;
        J.capacity = M.capacity;
        J.thresh = M.thresh;
        J.H = wh;
        J.W = ww;
        J.tpitch = (M.d_to_edgemap || M.d_to_orientation) ? M.tpitch : M.W;
        J.ntx = (ww + TW - 1) / TW;
        J.nty = (wh + TH - 1) / TH;
        J.tile_begin = (int)tiles;
        tiles += (long long)J.ntx * J.nty;
        J.nchunks = (int)(((long long)wh * J.nw + CHUNKW - 1) / CHUNKW);
        J.chunk_begin = (int)chunks;
        chunks += J.nchunks;
        p->max_tiles = std::max(p->max_tiles, J.ntx * J.nty);
        p->max_chunks = std::max(p->max_chunks, J.nchunks);
    }
    if (tiles >= (1LL << 31) || chunks >= (1LL << 31)) { toed_subpix_destroy(p); return TOED_ERR_ARG; }
    p->ntiles = tiles;
    p->nchunks = chunks;
    if (hipMalloc(&p->d_chunk, sizeof(int) * (size_t)chunks) != hipSuccess ||
        hipMalloc(&p->d_jobs, sizeof(SpJob) * (size_t)n) != hipSuccess) {
        toed_subpix_destroy(p);
        return TOED_ERR_ALLOC;
    }
    if (hipMemcpy(p->d_jobs, p->jobs.data(), sizeof(SpJob) * n, hipMemcpyHostToDevice) != hipSuccess) {
        toed_subpix_destroy(p);
        return TOED_ERR_HIP;
    }
    *out = p;
    return TOED_OK;
}

int toed_subpix_run(toed_subpix *p, int *d_counts, int *d_flags) {
    if (!p) return TOED_ERR_ARG;
    SPCHK(hipSetDevice(p->device));
    bool changed = false;
    for (int m = 0; m < p->n; ++m) {
        int *c = d_counts ? d_counts + m : nullptr, *f = d_flags ? d_flags + m : nullptr;
        if (p->jobs[m].count_out != c || p->jobs[m].flags_out != f) changed = true;
        p->jobs[m].count_out = c;
        p->jobs[m].flags_out = f;
    }
    hipStream_t s = (hipStream_t)toed_get_stream(p->ctx);
    if (changed || p->dirty)
        SPCHK(hipMemcpyAsync(p->d_jobs, p->jobs.data(), sizeof(SpJob) * p->n, hipMemcpyHostToDevice, s));
    p->dirty = false;
    hipLaunchKernelGGL(sp_reset_kernel, dim3((p->max_tiles + 4 * 256 - 1) / (4 * 256), p->n), dim3(256), 0, s,
                       p->d_jobs);
    if (p->timing) SPCHK(hipEventRecord(p->ev[0], s));
    if (p->slog2) hipLaunchKernelGGL(sp_resize_kernel, dim3(p->max_tiles, p->n), dim3(NT), 0, s, p->d_jobs, p->n);
    hipLaunchKernelGGL(sp_orient_nms_kernel, dim3(p->max_tiles, p->n), dim3(NT), 0, s, p->d_jobs, p->n);
    if (p->timing) SPCHK(hipEventRecord(p->ev[1], s));
    if (p->taps.K <= MAXK)
        hipLaunchKernelGGL(sp_isolate_kernel<MAXK>, dim3(p->max_tiles, p->n), dim3(NT), 0, s, p->d_jobs, p->taps);
    else
        hipLaunchKernelGGL(sp_isolate_kernel<MAXK_WIDE>, dim3(p->max_tiles, p->n), dim3(NT), 0, s, p->d_jobs,
                           p->taps);
    if (p->timing) SPCHK(hipEventRecord(p->ev[2], s));
    hipLaunchKernelGGL(sp_scan_chunks, dim3(p->max_chunks, p->n), dim3(256), 0, s, p->d_jobs, p->n, p->d_chunk);
    hipLaunchKernelGGL(sp_scan_maps, dim3(p->n), dim3(1024), 0, s, p->d_jobs, p->d_chunk);
    hipLaunchKernelGGL(sp_prefix_kernel, dim3(p->max_chunks, p->n), dim3(256), 0, s, p->d_jobs, p->n, p->d_chunk);
    if (p->timing) SPCHK(hipEventRecord(p->ev[3], s));
    const int st = launch_correct_r(s, p);
    if (st) return st;
    if (p->timing) SPCHK(hipEventRecord(p->ev[4], s));
    bool maps = false;
    for (const auto &J : p->jobs) maps = maps || J.to_e || J.to_o;
    if (maps) hipLaunchKernelGGL(sp_maps_kernel, dim3(p->max_tiles, p->n), dim3(NT), 0, s, p->d_jobs, p->n);
    if (p->timing) SPCHK(hipEventRecord(p->ev[5], s));
    SPCHK(hipGetLastError());
    p->grad_ran = p->grad;
    return TOED_OK;
}

int toed_subpix_set_maps(toed_subpix *p, const float *const *d_maps) {
    if (!p || !d_maps) return TOED_ERR_ARG;
    for (int m = 0; m < p->n; ++m)
        if (!d_maps[m]) return TOED_ERR_ARG;
    for (int m = 0; m < p->n; ++m) {
        if (p->slog2) p->jobs[m].src_map = d_maps[m];   // resized into the work plane at each run
        else p->jobs[m].map = d_maps[m];
    }
    p->dirty = true;
    return TOED_OK;
}

int toed_subpix_set_grad(toed_subpix *p, int on) {
    if (!p || (on && p->slog2)) return TOED_ERR_ARG;     // the adjoint covers interp_n = 0
    SPCHK(hipSetDevice(p->device));
    // per map: tq, dirs (H x W x 16 B), gth (H x W x 4 B), the edgel
    // count and the edgel records (min(capacity, H W) x (128 + 64) B: an edgel
    // per pixel at most)
    auto rec_rows = [](const SpJob &J) { return (size_t)std::min<long long>(J.capacity, (long long)J.H * J.W); };
    if (on && !p->d_gwork) {
        size_t total = 0;
        p->goff.assign(p->n, 0);
        for (int m = 0; m < p->n; ++m) {
            const SpJob &J = p->jobs[m];
            if (J.H < 5 || J.W < 5) return TOED_ERR_ARG;   // single reflection of the smoothing
            const size_t px = (size_t)J.H * J.W;
            p->goff[m] = total;
            total += align_up(px * 16) * 2 + align_up(px * 4) + align_up(4) + align_up(rec_rows(J) * 128) +
                     align_up(rec_rows(J) * 64);
        }
        if (hipMalloc(&p->d_gwork, total) != hipSuccess) return TOED_ERR_ALLOC;
        if (hipMalloc(&p->d_grads, sizeof(SpGrad) * p->n) != hipSuccess) return TOED_ERR_ALLOC;
        SPCHK(hipMemset(p->d_gwork, 0, total));
    }
    for (int m = 0; m < p->n; ++m) {
        SpJob &J = p->jobs[m];
        J.tq = nullptr; J.dirs = nullptr; J.rec = nullptr; J.brec = nullptr; J.gcount = nullptr;
        J.gth = nullptr;
        if (!on) continue;
        const size_t px = (size_t)J.H * J.W;
        char *b = p->d_gwork + p->goff[m];
        J.tq = (float4 *)b; b += align_up(px * 16);
        J.dirs = (double2 *)b; b += align_up(px * 16);
        J.gth = (float *)b; b += align_up(px * 4);
        J.gcount = (int *)b; b += align_up(4);
        J.rec = (float4 *)b; b += align_up(rec_rows(J) * 128);
        J.brec = (float4 *)b;
        // the edgel kernel's grid: enough workgroups to fill the device over
        // all maps, none beyond the largest map's records
        const long long need = ((long long)rec_rows(J) + NT - 1) / NT;
        p->eblocks = (int)std::max<long long>(p->eblocks, std::min<long long>(need, std::max(64, 4096 / p->n)));
    }
    p->grad = on != 0;
    p->grad_ran = false;
    SPCHK(hipMemcpy(p->d_jobs, p->jobs.data(), sizeof(SpJob) * p->n, hipMemcpyHostToDevice));
    return TOED_OK;
}

int toed_subpix_backward(toed_subpix *p, const toed_subpix_grad_t *items) {
    if (!p || !items) return TOED_ERR_ARG;
    if (!p->grad_ran) return TOED_ERR_ARG;   // the last run must have had the differentiable form on
    SPCHK(hipSetDevice(p->device));
    std::vector<SpGrad> g(p->n);
    for (int m = 0; m < p->n; ++m) {
        if (!items[m].d_gmap || items[m].gpitch < p->jobs[m].W) return TOED_ERR_ARG;
        if (!items[m].d_gout && p->jobs[m].capacity > 0) return TOED_ERR_ARG;
        g[m].gout = (const float4 *)items[m].d_gout;
        g[m].gmap = items[m].d_gmap;
        g[m].gpitch = items[m].gpitch;
        g[m].pad = 0;
    }
    hipStream_t s = (hipStream_t)toed_get_stream(p->ctx);
    SPCHK(hipMemcpyAsync(p->d_grads, g.data(), sizeof(SpGrad) * p->n, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(sp_bw_clear_kernel, dim3(256, p->n), dim3(256), 0, s, p->d_jobs);
    if (p->eblocks > 0)
        hipLaunchKernelGGL(sp_bw_edgel_kernel, dim3(p->eblocks, p->n), dim3(NT), 0, s, p->d_jobs, p->d_grads);
    const int st = launch_bw_tile_r(s, p);
    if (st) return st;
    hipLaunchKernelGGL(sp_bw_dense_kernel, dim3(p->max_tiles, p->n), dim3(NT), 0, s, p->d_jobs, p->d_grads,
                       p->taps.R);
    SPCHK(hipGetLastError());
    return TOED_OK;
}

int toed_subpix_set_timing(toed_subpix *p, int on) {
    if (!p) return TOED_ERR_ARG;
    SPCHK(hipSetDevice(p->device));
    if (on && !p->ev[0])
        for (auto &e : p->ev) SPCHK(hipEventCreate(&e));
    p->timing = on != 0;
    return TOED_OK;
}

int toed_subpix_last_times(toed_subpix *p, float *ms_orient_nms, float *ms_correct, float *ms_list) {
    if (!p || !p->ev[0]) return TOED_ERR_ARG;
    SPCHK(hipEventSynchronize(p->ev[5]));
    // orientation + NMS; isolation + correction; scans, ranks and maps
    float t[5] = {0.f, 0.f, 0.f, 0.f, 0.f};
    for (int k = 0; k < 5; ++k) SPCHK(hipEventElapsedTime(&t[k], p->ev[k], p->ev[k + 1]));
    const float a = t[0], b = t[1] + t[3], c = t[2] + t[4];
    if (ms_orient_nms) *ms_orient_nms = a;
    if (ms_correct) *ms_correct = b;
    if (ms_list) *ms_list = c;
    return TOED_OK;
}

int toed_subpix_destroy(toed_subpix *p) {
    if (!p) return TOED_ERR_ARG;
    (void)hipSetDevice(p->device);
    (void)hipStreamSynchronize((hipStream_t)toed_get_stream(p->ctx));
    (void)hipFree(p->d_work);
    (void)hipFree(p->d_chunk);
    (void)hipFree(p->d_nsurv);
    (void)hipFree(p->d_jobs);
    for (auto &e : p->ev)
        if (e) (void)hipEventDestroy(e);
    (void)hipFree(p->d_gwork);
    (void)hipFree(p->d_grads);
    delete p;
    return TOED_OK;
}

}  // extern "C"
