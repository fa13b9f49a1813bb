// TOED detection path for MI355X (gfx950): fused tile kernel, segmented scan,
// raster-order scatter, host orchestration and the C ABI (toed_capi.h).
//
// Algorithm (restated from the reference's CPU detector, cpu_toed.cpp:128-545):
//   * four interp samples per original pixel (i, j): (2i+a, 2j+b), a, b in {0,1},
//     each a Gaussian-derivative response with x shifted by b/2 and y by a/2;
//     phase (0,0) uses the R0 = ceil(4 sigma) window, the others R1 = R0 + 1
//     (cpu_toed.cpp:205-344);
//   * |grad I| threshold, octant NMS with linear interpolation of the magnitude
//     one step along the gradient, Fisher-Naidu parabola (cpu_toed.cpp:400-494);
//   * third-order orientation from the nine responses (cpu_toed.cpp:227-232);
//   * raster-ordered N x 4 list [x, y, theta, strength] (cpu_toed.cpp:510-538).
//
// MI355X design:
//   K1 toed_tiles_kernel   one workgroup (4 wavefronts) per TH x 32 tile of
//        original pixels = 2TH x 64 interp pixels.  The image tile plus halo is
//        staged once in LDS; the 2-D filters are applied separably (vertical
//        pass, then horizontal pass) in FP32 in three rounds that reuse two
//        LDS planes:
//        (R0-window G, G') -> phase (0,0); + end taps -> phase (1,0);
//        (shifted G, G') -> phases (0,1), (1,1).  Ix, Iy of the tile plus a
//        one-pixel interp halo stay in LDS; each wavefront runs NMS on one
//        64-wide interp row at a time, decides near-peaks and the parabola in
//        FP64, and marks survivors in per-row masks: per-(row, tile) counts go to
//        HBM, records (position, strength) and raster keys to a per-tile,
//        per-phase staging area; a tile with edgels appends itself to a list.
//   K2 toed_scan_chunks / toed_scan_images   exclusive scan of the counts
//        (row-major over (interp row, tile column)), segmented per image.
//   K2b toed_qad_kernel   orientation mode 1 only: the orientation evaluated at
//        each edgel's subpixel location instead (one wavefront per edgel).
//   K3 toed_finalize_kernel   a fixed grid walking the list of tiles with
//        edgels: the nine responses of each staged edgel in FP64 from the
//        tile's image region (8 lanes per edgel), its third-order orientation,
//        and the edgel to offset[row, tile] + rank (raster order).
// No MFMA: the filters are short separable stencils, not a contraction.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <type_traits>
#include <vector>

#include "toed_capi.h"
#include "toed_adjoint.hpp"

namespace toed {

constexpr int TW = 32;              // tile width in original pixels -> 64 interp columns
constexpr int NT = 256;             // threads per tile workgroup (4 wavefronts)
constexpr int MAXR1 = 9;            // largest supported R1 (sigma <= 2)
constexpr int MAXTAPS = 2 * MAXR1 + 1;
constexpr int CHUNK = 1024;         // scan chunk (entries)
constexpr int NGROUPS = 3;          // supported window radii R1 = 5, 7, 9 (sigma 1, 1.5, 2)
constexpr int BW_CHUNK = 64;        // edgels per pre-pass chunk of the tile backward
constexpr int BW_G = 8;             // lanes per edgel in the tile backward

// Filter taps for one sigma.  Families: 0 G, 1 G', 2 G'', 3 G''' sampled at
// x = k - R1 (unshifted) and x = k - R1 + 1/2 (shifted), k = 0..2R1
// (closed forms of cpu_toed.cpp:141-162, window rule of cpu_toed.cpp:131-132).
struct TapSet {
    int R0, R1, pad0, pad1;
    float fu[2][MAXTAPS];
    float fs[2][MAXTAPS];
    double du[4][MAXTAPS];
    double ds[4][MAXTAPS];
    double dsu[4][MAXTAPS];     // d/dsigma of du (window radii held fixed)
    double dss[4][MAXTAPS];     // d/dsigma of ds
};

// Per-image outputs of a backward pass over a batch plan.
struct GradOut {
    const float *gout;         // dL/d[x, y, theta, strength] per output row
    float *gimg;               // dL/dimage (accumulated)
    double *gsigma;            // dL/dsigma (accumulated), may be null
    int gpitch, pad;
};

// One image of a launch.  All offsets are in elements of their arrays.
struct ImageJob {
    const void *img;
    float *out;
    int *count_out;
    int *flags_out;
    const TapSet *taps;
    long long capacity;
    long long counts_base;     // first entry in counts/offsets (chunk aligned)
    long long chunk_base;      // first scan chunk
    long long stage_base;      // first staging record
    int in_type, H, W, pitch;
    int R0, R1, ntc, ntr;
    int tile_begin, nchunks, border, stage_cap;   // stage_cap = records per tile and phase
    float thresh;
    int qad;                   // 1: orientation at the subpixel location (toed_set_orientation_mode)
    int *index_out;            // optional interp index per edgel
    long long *cbase;          // compact output: the device slot holding this image's first row (else null)
};

// Survivors of a tile, per phase: the finished edgel [x, y, theta, strength]
// and its raster key (local row | local col << 8 | rank-in-row << 16).
struct StageBuf {
    float4 *rec;
    uint32_t *key;
};

__host__ __device__ inline int ceil_div(int a, int b) { return (a + b - 1) / b; }

// Read-only tables (jobs, taps) are read through the constant address space so
// that wavefront-uniform reads become scalar loads into SGPRs.
#define TOED_CONST __attribute__((address_space(4)))
template <typename T> __device__ __forceinline__ const TOED_CONST T *as_const(const T *p) {
    return (const TOED_CONST T *)p;
}

// Index of the job owning a tile (0 for single images).
__device__ __forceinline__ int job_index(const ImageJob *jobs, int njobs, int tile) {
    if (!jobs) return 0;
#if defined(__HIP_DEVICE_COMPILE__)
    const TOED_CONST ImageJob *cj = as_const(jobs);
    int lo = 0, hi = njobs - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (cj[mid].tile_begin <= tile) lo = mid; else hi = mid - 1;
    }
    return lo;
#else
    (void)njobs; (void)tile;
    return 0;
#endif
}

// The job of a tile: the kernel-argument job for single images, else the table
// entry found by binary search over tile_begin (uniform across the workgroup).
__device__ __forceinline__ ImageJob load_job(const ImageJob *jobs, int njobs, const ImageJob &job0, int tile) {
    if (!jobs) return job0;
#if defined(__HIP_DEVICE_COMPILE__)
    const TOED_CONST ImageJob *cj = as_const(jobs);
    int lo = 0, hi = njobs - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (cj[mid].tile_begin <= tile) lo = mid; else hi = mid - 1;
    }
    return cj[lo];
#else
    return job0;
#endif
}

__device__ inline const ImageJob &job_for_tile(const ImageJob *jobs, int njobs, int tile) {
    int lo = 0, hi = njobs - 1;
    while (lo < hi) {   // last job with tile_begin <= tile
        int mid = (lo + hi + 1) >> 1;
        if (jobs[mid].tile_begin <= tile) lo = mid; else hi = mid - 1;
    }
    return jobs[lo];
}

__device__ inline int job_for_chunk(const ImageJob *jobs, int njobs, long long chunk) {
    int lo = 0, hi = njobs - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (jobs[mid].chunk_base <= chunk) lo = mid; else hi = mid - 1;
    }
    return lo;
}

template <typename T> __device__ inline float load_px(const void *p, long long idx) {
    return (float)((const T *)p)[idx];
}

// ---------------------------------------------------------------------------
// LDS geometry of one tile
// ---------------------------------------------------------------------------
// EB: bytes per staged image pixel, 1 for u8 images (the pixels are exact in
// a byte and are widened as the vertical pass reads them: the image region
// shrinks by three quarters, which lets a fifth workgroup onto a CU) and 4
// for f32 images.
template <int R1, int TH, int EB = 4> struct Geo {
    static_assert(EB == 1 || EB == 4, "staged pixel size");
    static constexpr int R0 = R1 - 1;
    static constexpr int NRI = TH + 2 + 2 * R1;            // staged image rows
    static constexpr int NCI = TW + 2 + 2 * R1;            // staged image cols
    static constexpr int LDI = NCI | 1;                    // odd pitch (floats): conflict-free columns
    // u8 rows: the dword-aligned global columns from gj0 & ~3 land on LDS
    // dwords, so image column c sits at byte SHB + c of its row; gj0 =
    // 32 tc - 1 - R1, so SHB = gj0 & 3 is the same for every tile
    static constexpr int SHB = (-1 - R1) & 3;
    static constexpr int LDB = 4 * (((SHB + NCI + 3) / 4) | 1);    // row pitch in bytes (odd in dwords)
    static constexpr int IMG_FLOATS = EB == 4 ? NRI * LDI : NRI * LDB / 4;
    static constexpr int NVR = TH + 2;                     // vertical-pass output rows
    static constexpr int RH = 2 * TH + 2, RW = 2 * TW + 2; // interp results incl. 1-px halo
    static constexpr int LDR = RW;
    // LDS layout (floats): staged image | Ix, Iy with 1-px halo | two
    // vertical-pass planes (NVR rows x NCI cols, pitch LDP), later reused for the
    // near-peak queue (which becomes the survivor queue) | row masks.
    // |grad I| is recomputed from Ix, Iy where needed (no third plane)
    static constexpr int OFF_IMG = 0;
    static constexpr int OFF_IX = (IMG_FLOATS + 1) & ~1;
    static constexpr int OFF_IY = OFF_IX + RH * LDR;
    static constexpr int OFF_P0 = (OFF_IY + RH * LDR + 1) & ~1;      // 8-byte aligned
    // vertical-pass plane pitch: the smallest >= NCI that is 1 mod 32, so that
    // the horizontal pass's window loads (lanes along a row 3 columns apart,
    // the next row 11 lanes on) hit 32 distinct banks in each half-wavefront
    static constexpr int LDP = (NCI + 30) / 32 * 32 + 1;
    static constexpr int OFF_P1 = OFF_P0 + NVR * LDP;
    static constexpr int QMAX = 2 * TH * 2 * TW;                     // survivor queue (ints)
    static constexpr int P_END = OFF_P1 + NVR * LDP;
    static constexpr int Q_END = OFF_P0 + QMAX;
    static constexpr int OFF_CAND = OFF_P0;      // near-peak queue (ints), then the survivor queue
    static constexpr int OFF_MASK = ((P_END > Q_END ? P_END : Q_END) + 1) & ~1;   // 2TH u64 row masks
    static constexpr int TOTAL = OFF_MASK + 2 * (2 * TH);
    // Work items.  Vertical pass: (col, block of KV rows), the NVR rows exactly.
    // Horizontal pass, one phase at a time: (row, block of KH cols) over the
    // HR = TH + 1 rows and HC = TW + 1 columns whose results land inside the
    // interp planes (17 x 33 at TH = 16: no output outside them is computed).
    // Both stay within three wavefronts (the fourth waits at the barrier),
    // with a long FMA run per loaded value.
    static constexpr int KV = TH >= 16 ? 6 : 5;
    static constexpr int NBV = (NVR + KV - 1) / KV;
    static constexpr int KH = 3;
    static constexpr int HR = TH + 1, HC = TW + 1;
    static constexpr int NBH = (HC + KH - 1) / KH;
    static_assert(NBV * KV == NVR && NBH * KH == HC, "blocks cover the pass outputs exactly");
    static_assert(NCI * NBV <= 3 * 64 && HR * NBH <= 3 * 64, "three wavefronts per pass");
    // every read stays inside its array: the vertical pass reads image rows
    // up to NVR - 1 + 2 R1 = NRI - 1, the horizontal pass plane rows up to
    // HR = NVR - 1 and columns up to HC + 2 R1 = NCI - 1
};

template <typename IN_T> constexpr int stage_bytes() { return std::is_same<IN_T, uint8_t>::value ? 1 : 4; }


// The 2-D filters run vertical pass first: a 16 x 32 tile is wider than it
// is tall, so the first pass's halo is smaller that way round
// ((TH + 2) x (TW + 2 + 2 R1) first-pass outputs instead of
// (TH + 2 + 2 R1) x (TW + 2)), 12% fewer FMAs at sigma 2.
//
// Vertical pass over the staged image columns into the two planes (P0: G in
// y, P1: G' in y).  mode 0: R0-window taps (fresh); mode 1: add the two end
// taps (R0 -> R1 window); mode 2: shifted taps (fresh).
template <int R1, int TH, int MODE, int EB>
__device__ __forceinline__ void vpass(float *lds, const TOED_CONST float *t0, const TOED_CONST float *t1) {
    using Gm = Geo<R1, TH, EB>;
    float *P0 = lds + Gm::OFF_P0;
    float *P1 = lds + Gm::OFF_P1;
    constexpr int R = (MODE == 0) ? Gm::R0 : R1;
    constexpr int KV = Gm::KV;
    const int it = threadIdx.x;
    if (it >= Gm::NCI * Gm::NBV) return;
    const int u = it % Gm::NCI;
    const int v0 = (it / Gm::NCI) * KV;
    // image row v0 + t, column u
    const float *colf = lds + Gm::OFF_IMG + v0 * Gm::LDI + u;
    const uint8_t *colb = (const uint8_t *)(lds + Gm::OFF_IMG) + v0 * Gm::LDB + Gm::SHB + u;
    auto px = [&](int t) -> float {
        if constexpr (EB == 4) return colf[t * Gm::LDI];
        else return (float)colb[t * Gm::LDB];
    };
    if (MODE == 1) {
#pragma unroll
        for (int d = 0; d < KV; ++d) {
            const float lo = px(d), hi = px(d + 2 * R1);
            P0[(v0 + d) * Gm::LDP + u] += hi * t0[0] + lo * t0[2 * R1];
            P1[(v0 + d) * Gm::LDP + u] += hi * t1[0] + lo * t1[2 * R1];
        }
    } else {
        float w[KV + 2 * R1];
#pragma unroll
        for (int t = 0; t < KV + 2 * R1; ++t) w[t] = px(t);
#pragma unroll
        for (int d = 0; d < KV; ++d) {
            // the first term as a product (an fma onto +0 would cost one more op)
            float a0 = w[d + R1 + R] * t0[R1 - R], a1 = w[d + R1 + R] * t1[R1 - R];
#pragma unroll
            for (int q = -R + 1; q <= R; ++q) {
                const float v = w[d + R1 - q];
                a0 = fmaf(v, t0[q + R1], a0);
                a1 = fmaf(v, t1[q + R1], a1);
            }
            P0[(v0 + d) * Gm::LDP + u] = a0;
            P1[(v0 + d) * Gm::LDP + u] = a1;
        }
    }
}

// Horizontal pass for the phase with y shift A and x shift B (0/1) from the
// current planes: Ix = sum_x P0 * X1, Iy = sum_x P1 * X0 (x taps X0, X1) at
// interp row 2v - 1 + A, column 2u - 1 + B, for the plane rows v = 1 - A ..
// TH + 1 - A and columns u = 1 - B .. TW + 1 - B: exactly the results inside
// the interp planes.  Returns the largest Ix^2 + Iy^2 among the thread's
// results (0 for a thread without an item).
template <int R1, int TH, int R, int A, int B, int EB>
__device__ __forceinline__ float hpass(float *lds, const TOED_CONST float *x0, const TOED_CONST float *x1) {
    using Gm = Geo<R1, TH, EB>;
    float *RX = lds + Gm::OFF_IX;
    float *RY = lds + Gm::OFF_IY;
    constexpr int KH = Gm::KH;
    constexpr int NW = KH + 2 * R1;
    const int it = threadIdx.x;
    if (it >= Gm::HR * Gm::NBH) return 0.f;
    // column blocks fastest: 11 lanes per plane row
    const int v = (1 - A) + it / Gm::NBH;
    const int u0 = (1 - B) + (it % Gm::NBH) * KH;
    // the windows' first elements as opaque registers: the loads then take
    // their small constant offsets from one base per plane (a visible constant
    // base makes the compiler add each large plane offset separately)
    int base0 = Gm::OFF_P0 + v * Gm::LDP + u0, base1 = base0 + (Gm::OFF_P1 - Gm::OFF_P0);
    asm volatile("" : "+v"(base0), "+v"(base1));
    const float *p0 = lds + base0;
    const float *p1 = lds + base1;
    float w0[NW], w1[NW];
#pragma unroll
    for (int t = 0; t < NW; ++t) {
        w0[t] = p0[t];
        w1[t] = p1[t];
    }
    const int rr = 2 * v - 1 + A;
    float mag2 = 0.f;
#pragma unroll
    for (int d = 0; d < KH; ++d) {
        float ix = w0[d + R1 + R] * x1[R1 - R], iy = w1[d + R1 + R] * x0[R1 - R];
#pragma unroll
        for (int q = -R + 1; q <= R; ++q) {
            ix = fmaf(w0[d + R1 - q], x1[q + R1], ix);
            iy = fmaf(w1[d + R1 - q], x0[q + R1], iy);
        }
        const int cc = 2 * (u0 + d) - 1 + B;
        RX[rr * Gm::LDR + cc] = ix;
        RY[rr * Gm::LDR + cc] = iy;
        mag2 = fmaxf(mag2, ix * ix + iy * iy);
    }
    return mag2;
}

// Octant choice and neighbour interpolation of cpu_toed.cpp:415-462 for a pixel
// whose magnitude plane is reachable through M(dr, dc).  Returns 0 for a NaN
// gradient (no branch applies).  T is float (selection) or double (values).
template <typename T, typename MF>
__device__ __forceinline__ int octant(T gx, T gy, T nx, T ny, MF M, T &slope, T &fp, T &fm) {
    if (gx >= 0 && gy >= 0) {
        if (gx >= gy) { slope = ny / nx; fp = M(0, 1) * (1 - slope) + M(1, 1) * slope;  fm = M(0, -1) * (1 - slope) + M(-1, -1) * slope; }
        else          { slope = nx / ny; fp = M(1, 0) * (1 - slope) + M(1, 1) * slope;  fm = M(-1, 0) * (1 - slope) + M(-1, -1) * slope; }
    } else if (gx < 0 && gy >= 0) {
        if (fabs(gx) < gy) { slope = -nx / ny; fp = M(1, 0) * (1 - slope) + M(1, -1) * slope; fm = M(-1, 0) * (1 - slope) + M(-1, 1) * slope; }
        else               { slope = -ny / nx; fp = M(0, -1) * (1 - slope) + M(1, -1) * slope; fm = M(0, 1) * (1 - slope) + M(-1, 1) * slope; }
    } else if (gx < 0 && gy < 0) {
        if (fabs(gx) >= fabs(gy)) { slope = ny / nx; fp = M(0, -1) * (1 - slope) + M(-1, -1) * slope; fm = M(0, 1) * (1 - slope) + M(1, 1) * slope; }
        else                      { slope = nx / ny; fp = M(-1, 0) * (1 - slope) + M(-1, -1) * slope; fm = M(1, 0) * (1 - slope) + M(1, 1) * slope; }
    } else if (gx >= 0 && gy < 0) {
        if (gx < fabs(gy)) { slope = -nx / ny; fp = M(-1, 0) * (1 - slope) + M(-1, 1) * slope; fm = M(1, 0) * (1 - slope) + M(1, -1) * slope; }
        else               { slope = -ny / nx; fp = M(0, 1) * (1 - slope) + M(-1, 1) * slope; fm = M(0, -1) * (1 - slope) + M(1, -1) * slope; }
    } else {
        return 0;
    }
    return 1;
}

// NMS over the tile's 2TH x 64 interp pixels (cpu_toed.cpp:400-494), in
// three steps so that the FP64 work runs on compacted lanes:
//  1. dense FP32 filter, one 64-wide row per wavefront step: threshold, then
//     the peak test against the magnitudes interpolated one step along the
//     gradient, with a 1e-5 relative slack and a fast reciprocal for the slope;
//     near-peaks are queued (ballot compaction);
//  2. every queued pixel re-decided exactly in FP64 on the FP32 magnitudes
//     (octant rules, peak test, Fisher-Naidu parabola, |s*| <= sqrt 2);
//     survivors take a staging slot of their phase, join the orientation
//     queue and set their bit in the row's 64-bit mask;
//  3. per-(row, tile) counts from the masks (the raster ranks are taken from
//     the same masks by the orientation pass).
template <int R1, int TH, int EB>
__device__ __forceinline__ void nms_tile(float *lds, const ImageJob &J, int tr, int tc, long long tile_local,
                                int *__restrict__ counts, StageBuf stg, int *sh_cnt, int *cand, int *queue,
                                unsigned long long *rowmask) {
    using Gm = Geo<R1, TH, EB>;
    const float *RX = lds + Gm::OFF_IX;
    const float *RY = lds + Gm::OFF_IY;
    // |grad I| from Ix, Iy (hardware square root, 1 ulp) wherever it is needed
    auto MG = [&](int o) -> float {
        const float x = RX[o], y = RY[o];
        return __builtin_amdgcn_sqrtf(x * x + y * y);
    };
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int IH = 2 * J.H, IW = 2 * J.W;
    const int R0g = 2 * tr * TH, C0g = 2 * tc * TW;
    const unsigned long long lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    // step 1: FP32 filter over the wavefront's RPW rows r = wave + 4k (four
    // wavefronts): the
    // centre values of all its rows are loaded up front (their latencies
    // overlap), each row's near-peaks are balloted, and the wavefront takes
    // its queue slots with one LDS atomic for all its rows
    constexpr int RPW = 2 * TH / (NT / 64);
    float gxs[RPW], gys[RPW];
#pragma unroll
    for (int k = 0; k < RPW; ++k) {
        const int o = (wave + (NT / 64) * k + 1) * Gm::LDR + lane + 1;
        gxs[k] = RX[o];
        gys[k] = RY[o];
    }
    unsigned long long bal[RPW];
    int tot = 0;
#pragma unroll
    for (int k = 0; k < RPW; ++k) {
        const int r = wave + (NT / 64) * k;
        const int R = R0g + r, C = C0g + lane;
        const int rr = r + 1, cc = lane + 1;
        bool cnd = false;
        const bool inside = R >= J.border && R < IH - J.border && C >= J.border && C < IW - J.border;
        const float gx = gxs[k], gy = gys[k];
        const float f = __builtin_amdgcn_sqrtf(gx * gx + gy * gy);
        if (inside && f > J.thresh) {
            const float ax = fabsf(gx), ay = fabsf(gy);
            if (!(ax < 10e-6f && ay < 10e-6f)) {
                // x-major iff |gx| >= |gy|; step signs from the gradient signs
                const int sx = gx >= 0.f ? 1 : -1, sy = gy >= 0.f ? 1 : -1;
                const bool xm = ax >= ay;
                const float sl = (xm ? ay : ax) * __builtin_amdgcn_rcpf(xm ? ax : ay);
                const int pr = xm ? 0 : sy, pc = xm ? sx : 0;        // primary neighbour
                const float mp = MG((rr + pr) * Gm::LDR + cc + pc), md = MG((rr + sy) * Gm::LDR + cc + sx);
                const float np = MG((rr - pr) * Gm::LDR + cc - pc), nd = MG((rr - sy) * Gm::LDR + cc - sx);
                const float fp = mp + sl * (md - mp), fm = np + sl * (nd - np);
                cnd = f >= fmaxf(fp, fm) * (1.0f - 1e-5f);
            }
        }
        bal[k] = __ballot(cnd);
        tot += __popcll(bal[k]);
    }
    if (tot) {
        int b0 = 0;
        if (lane == 0) b0 = atomicAdd(&sh_cnt[5], tot);
        b0 = __shfl(b0, 0);
#pragma unroll
        for (int k = 0; k < RPW; ++k) {
            if ((bal[k] >> lane) & 1ull) cand[b0 + __popcll(bal[k] & lt)] = (wave + (NT / 64) * k) | (lane << 6);
            b0 += __popcll(bal[k]);
        }
    }
    __syncthreads();
    // step 2: exact FP64 decision and parabola on the queued pixels
    const int ncand = sh_cnt[5];
    const double SQ2 = 1.4142135623730951;
    // the queue of survivors is written over the candidates: each thread reads
    // its candidates first
    constexpr int CPT = Gm::QMAX / NT;
    int qv[CPT];
#pragma unroll
    for (int j = 0; j < CPT; ++j) {
        const int k = threadIdx.x + j * NT;
        qv[j] = k < ncand ? cand[k] : -1;
    }
    __syncthreads();
#pragma unroll 1
    for (int j = 0; j < CPT; ++j) {
        const int q = qv[j];
        if (q < 0) break;
        const int r = q & 63, c = q >> 6;
        const int rr = r + 1, cc = c + 1;
        const double fd = MG(rr * Gm::LDR + cc);
        const double gxd = RX[rr * Gm::LDR + cc], gyd = RY[rr * Gm::LDR + cc];
        const double nxd = gxd / fd, nyd = gyd / fd;
        auto Md = [&](int dr, int dc) { return (double)MG((rr + dr) * Gm::LDR + cc + dc); };
        double sld, fpd, fmd;
        if (!octant<double>(gxd, gyd, nxd, nyd, Md, sld, fpd, fmd)) continue;
        if (!((fd > fmd && fd > fpd) || (fd > fmd && fd >= fpd) || (fd >= fmd && fd > fpd))) continue;
        const double s = sqrt(1.0 + sld * sld);
        const double A = (fmd + fpd - 2.0 * fd) / (2.0 * s * s);
        const double Bc = (fpd - fmd) / (2.0 * s);
        const double ss = -Bc / (2.0 * A);
        const double maxf = A * ss * ss + Bc * ss + fd;
        if (!(fabs(ss) <= SQ2)) continue;
        const double sgx = maxf * nxd, sgy = maxf * nyd;
        const int R = R0g + r, C = C0g + c;
        const int phase = 2 * (r & 1) + (c & 1);
        atomicOr(&rowmask[r], 1ull << c);
        const int idx = atomicAdd(&sh_cnt[phase], 1);
        if (idx < J.stage_cap) {
            const int slot = phase * J.stage_cap + idx;
            const long long g = J.stage_base + tile_local * 4 * J.stage_cap + slot;
            stg.rec[g] = make_float4((float)(((double)C + ss * nxd - 1.0) * 0.5),
                                     (float)(((double)R + ss * nyd - 1.0) * 0.5), 0.f,
                                     (float)sqrt(sgx * sgx + sgy * sgy));
            queue[atomicAdd(&sh_cnt[4], 1)] = r | (c << 6) | (slot << 12);
        }
    }
    __syncthreads();
    // step 3: survivors per (interp row, tile column)
    for (int r = threadIdx.x; r < 2 * TH; r += NT) {
        const int R = R0g + r;
        if (R < IH) counts[J.counts_base + (long long)R * J.ntc + tc] = __popcll(rowmask[r]);
    }
}

// Angle of the third-order vector (TOx, TOy): cpu_toed.cpp:227-232 divides it
// by its norm and takes atan2(TOx, -TOy).  A positive scale leaves the angle
// unchanged, so the FP64 vector is scaled exactly by a power of two (its
// larger component to [0.5, 1)) and the angle is taken in FP32, the precision
// it is stored in; a zero vector gives NaN, as the reference's 0/0 does.
__device__ __forceinline__ float to_angle(double tx, double ty) {
    const double m = fmax(fabs(tx), fabs(ty));
    if (!(m > 0.0)) return __builtin_nanf("");
    int e;
    (void)frexp(m, &e);
    return atan2f((float)ldexp(tx, -e), (float)ldexp(-ty, -e));
}

// Third-order orientation (cpu_toed.cpp:227-232) of staged edgels, in FP64,
// run by the scatter kernel on the tile's image region, which the workgroup
// has staged in LDS as FP64 (the window sums then read no conversions; rows
// and columns from 1 + R1 before the tile, zero outside
// the image, cpu_toed.cpp:207-208; row pitch FLDI): 8 lanes per edgel, lane l
// the window rows p = -R1 + l + 8k (the fewest idle lanes for 2R1+1 = 11, 15,
// 19 rows), the row sums for the four x families over the columns (a column
// per iteration, not unrolled: measured fastest), then the nine products with
// the row's y taps, summed across the 8 lanes
// (DPP).  Phase (0,0) uses the R0 window through taps whose end values are
// zero.  tapd: [family][tap][f] with families 0 unshifted, 1 unshifted with
// zero end taps, 2 shifted.  (ci, cj): the edgel's image pixel in the region.
constexpr int FLDI = (TW + 2 + 2 * MAXR1) | 1;            // region row pitch (elements)
constexpr int FROWS = 16 + 2 + 2 * MAXR1;                 // region rows at most (16-row tiles)

template <int R1, int GT>
__device__ __forceinline__ double theta_group(const double *img, int ci, int cj, int a, int b, int l,
                                              const double (*tapd)[MAXTAPS][4]) {
    constexpr int N = 2 * R1 + 1;
    constexpr int NK = (N + GT - 1) / GT;
    const int phase = 2 * a + b;
    const double(*X)[4] = tapd[b ? 2 : (phase == 0 ? 1 : 0)];
    const double(*Y)[4] = tapd[a ? 2 : (phase == 0 ? 1 : 0)];
    double acc[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    // row sums of the lane's rows (rows beyond the window read a valid row and
    // are dropped below), each column's taps read once for all of them
    double h[NK][4];
    const double *rows[NK];
#pragma unroll
    for (int k = 0; k < NK; ++k) {
        const int p = min(-R1 + l + GT * k, R1);
        rows[k] = img + (ci - p) * FLDI + cj;
        h[k][0] = h[k][1] = h[k][2] = h[k][3] = 0.0;
    }
#pragma unroll 1
    for (int qq = -R1; qq <= R1; ++qq) {
        const double *x = X[qq + R1];
        const double x0 = x[0], x1 = x[1], x2 = x[2], x3 = x[3];
#pragma unroll
        for (int k = 0; k < NK; ++k) {
            const double v = rows[k][-qq];
            h[k][0] = fma(v, x0, h[k][0]);
            h[k][1] = fma(v, x1, h[k][1]);
            h[k][2] = fma(v, x2, h[k][2]);
            h[k][3] = fma(v, x3, h[k][3]);
        }
    }
#pragma unroll
    for (int k = 0; k < NK; ++k) {
        const int p = -R1 + l + GT * k;
        if (p <= R1) {
            const double h0 = h[k][0], h1 = h[k][1], h2 = h[k][2], h3 = h[k][3];
            const double *y = Y[p + R1];
            const double y0 = y[0], y1 = y[1], y2 = y[2], y3 = y[3];
            acc[0] = fma(h1, y0, acc[0]);   // fx
            acc[1] = fma(h0, y1, acc[1]);   // fy
            acc[2] = fma(h2, y0, acc[2]);   // fxx
            acc[3] = fma(h1, y1, acc[3]);   // fxy
            acc[4] = fma(h0, y2, acc[4]);   // fyy
            acc[5] = fma(h2, y1, acc[5]);   // fxxy
            acc[6] = fma(h1, y2, acc[6]);   // fxyy
            acc[7] = fma(h3, y0, acc[7]);   // fxxx
            acc[8] = fma(h0, y3, acc[8]);   // fyyy
        }
    }
#pragma unroll
    for (int d = 0; d < 9; ++d) acc[d] = toeddpp::group_sum<GT>(acc[d]);
    const double fx = acc[0], fy = acc[1], fxx = acc[2], fxy = acc[3], fyy = acc[4];
    const double fxxy = acc[5], fxyy = acc[6], fxxx = acc[7], fyyy = acc[8];
    double tx = fx * (2 * fxx * fxx + 2 * fxy * fxy) + fy * (2 * fxx * fxy + 2 * fyy * fxy)
              + 2 * fx * fy * fxxy + fy * fy * fxyy + fx * fx * fxxx;
    double ty = fx * (2 * fxx * fxy + 2 * fyy * fxy) + fy * (2 * fyy * fyy + 2 * fxy * fxy)
              + 2 * fx * fy * fxyy + fx * fx * fxxy + fy * fy * fyyy;
    return (double)to_angle(tx, ty);
}

// Image staging of a tile (the tile + 1 + R1 halo; zero outside the image,
// cpu_toed.cpp:207-208): all of a thread's loads into registers first, then
// the LDS stores.  u8 images with 4-byte aligned rows move aligned dwords of
// 4 pixels (partial dwords at the image edges byte by byte).
template <int R1, int TH> struct StageRegs {
    using Gm = Geo<R1, TH>;
    static constexpr int NDW = Gm::NRI * ((Gm::NCI + 3) / 4 + 1);       // dwords per tile, at most
    static constexpr int NU = (NDW + NT - 1) / NT;
    uint32_t v[NU];
};

// The register route is taken for u8 images with 4-byte aligned rows (two
// dwords per thread); f32 tiles are staged pixel by pixel.
__device__ __forceinline__ bool stage_aligned_u8(const ImageJob &J) {
    return (J.pitch & 3) == 0 && ((uintptr_t)J.img & 3) == 0;
}

template <int R1, int TH>
__device__ __forceinline__ void stage_fetch_u8(const ImageJob &J, int tile, StageRegs<R1, TH> &sr) {
    using Gm = Geo<R1, TH>;
    const long long tile_local = tile - J.tile_begin;
    const int tr = (int)(tile_local / J.ntc), tc = (int)(tile_local % J.ntc);
    const int gi0 = tr * TH - 1 - R1, gj0 = tc * TW - 1 - R1;
    const uint8_t *base = (const uint8_t *)J.img;
    const int gc0 = gj0 & ~3;
    const int nd = (gj0 + Gm::NCI - gc0 + 3) >> 2;
#pragma unroll
    for (int u = 0; u < StageRegs<R1, TH>::NU; ++u) {
        const int e = threadIdx.x + u * NT;
        const int rr = e / nd, k = e - rr * nd;
        const int gi = gi0 + rr, col0 = gc0 + 4 * k;
        uint32_t w = 0;
        if (e < Gm::NRI * nd && gi >= 0 && gi < J.H) {
            const uint8_t *row = base + (long long)gi * J.pitch;
            if (col0 >= 0 && col0 + 3 < J.W) {
                w = *(const uint32_t *)(row + col0);
            } else {
#pragma unroll
                for (int b = 0; b < 4; ++b)
                    if (col0 + b >= 0 && col0 + b < J.W) w |= (uint32_t)row[col0 + b] << (8 * b);
            }
        }
        sr.v[u] = w;
    }
}

template <int R1, int TH>
__device__ __forceinline__ void stage_store_u8(float *lds, const ImageJob &J, int tile, const StageRegs<R1, TH> &sr) {
    // the fetched dwords are the LDS dwords of the byte rows (column c at byte SHB + c)
    using Gm = Geo<R1, TH, 1>;
    uint32_t *img = (uint32_t *)(lds + Gm::OFF_IMG);
    const long long tile_local = tile - J.tile_begin;
    const int tc = (int)(tile_local % J.ntc);
    const int gj0 = tc * TW - 1 - R1;
    const int gc0 = gj0 & ~3;
    const int nd = (gj0 + Gm::NCI - gc0 + 3) >> 2;
#pragma unroll
    for (int u = 0; u < StageRegs<R1, TH>::NU; ++u) {
        const int e = threadIdx.x + u * NT;
        if (e >= Gm::NRI * nd) continue;
        const int rr = e / nd, k = e - rr * nd;
        img[rr * (Gm::LDB / 4) + k] = sr.v[u];
    }
}

// f32 images and u8 images whose rows are not 4-byte aligned: pixel by pixel,
// straight to LDS
template <typename IN_T, int R1, int TH>
__device__ __forceinline__ void stage_direct(float *lds, const ImageJob &J, int tile) {
    constexpr int EB = stage_bytes<IN_T>();
    using Gm = Geo<R1, TH, EB>;
    float *img = lds + Gm::OFF_IMG;
    uint8_t *imgb = (uint8_t *)(lds + Gm::OFF_IMG) + Gm::SHB;
    const long long tile_local = tile - J.tile_begin;
    const int tr = (int)(tile_local / J.ntc), tc = (int)(tile_local % J.ntc);
    const int gi0 = tr * TH - 1 - R1, gj0 = tc * TW - 1 - R1;
    for (int e = threadIdx.x; e < Gm::NRI * Gm::NCI; e += NT) {
        const int rr = e / Gm::NCI, cc = e % Gm::NCI;
        const int gi = gi0 + rr, gj = gj0 + cc;
        const bool in = gi >= 0 && gi < J.H && gj >= 0 && gj < J.W;
        if constexpr (EB == 1) {
            imgb[rr * Gm::LDB + cc] = in ? ((const uint8_t *)J.img)[(long long)gi * J.pitch + gj] : (uint8_t)0;
        } else {
            img[rr * Gm::LDI + cc] = in ? load_px<IN_T>(J.img, (long long)gi * J.pitch + gj) : 0.f;
        }
    }
}

template <typename IN_T, int R1, int TH>
__device__ __forceinline__ void stage_tile(float *lds, const ImageJob &J, int tile) {
    if constexpr (std::is_same<IN_T, uint8_t>::value) {
        if (stage_aligned_u8(J)) {
            StageRegs<R1, TH> sr;
            stage_fetch_u8<R1, TH>(J, tile, sr);
            stage_store_u8<R1, TH>(lds, J, tile, sr);
            return;
        }
    }
    stage_direct<IN_T, R1, TH>(lds, J, tile);
}

// per-tile LDS counters: 4 phase counters, orientation-queue length,
// near-peak count; the survivor row masks
template <int R1, int TH, int EB> __device__ __forceinline__ void tile_setup(float *lds, int *sh_cnt) {
    using Gm = Geo<R1, TH, EB>;
    if (threadIdx.x < 6) sh_cnt[threadIdx.x] = 0;
    unsigned long long *rowmask = (unsigned long long *)(lds + Gm::OFF_MASK);
    if (threadIdx.x < 2 * TH) rowmask[threadIdx.x] = 0ull;
}

// Everything after staging: filters, NMS, orientation, counts.  The image and
// the counters of the tile are in LDS (a barrier precedes the call).
template <typename IN_T, int R1, int TH>
__device__ __forceinline__ void tile_compute(float *lds, int *sh_cnt, const ImageJob &J, int tile,
                          int *__restrict__ counts, StageBuf stg, int *__restrict__ tile_cnt, int *tlist) {
    constexpr int EB = stage_bytes<IN_T>();
    using Gm = Geo<R1, TH, EB>;
    const long long tile_local = tile - J.tile_begin;
    const int tr = (int)(tile_local / J.ntc), tc = (int)(tile_local % J.ntc);
    const TOED_CONST TapSet *tp = as_const(J.taps);
    unsigned long long *rowmask = (unsigned long long *)(lds + Gm::OFF_MASK);
    // round A: R0 window, unshifted -> phase (0,0)
    vpass<R1, TH, 0, EB>(lds, tp->fu[0], tp->fu[1]);
    __syncthreads();
    float mag2 = hpass<R1, TH, R1 - 1, 0, 0, EB>(lds, tp->fu[0], tp->fu[1]);
    __syncthreads();
    // round B: R1 window, unshifted in y -> phase (0,1) with shifted x taps
    vpass<R1, TH, 1, EB>(lds, tp->fu[0], tp->fu[1]);
    __syncthreads();
    mag2 = fmaxf(mag2, hpass<R1, TH, R1, 0, 1, EB>(lds, tp->fs[0], tp->fs[1]));
    __syncthreads();
    // round C: shifted in y -> phases (1,0) and (1,1), one x family at a time
    vpass<R1, TH, 2, EB>(lds, tp->fs[0], tp->fs[1]);
    __syncthreads();
    mag2 = fmaxf(mag2, hpass<R1, TH, R1, 1, 0, EB>(lds, tp->fu[0], tp->fu[1]));
    mag2 = fmaxf(mag2, hpass<R1, TH, R1, 1, 1, EB>(lds, tp->fs[0], tp->fs[1]));
    // A tile whose every Ix^2 + Iy^2 (halo included) is at most
    // thresh^2 (1 - 2^-20) has no NMS candidate: the test there is
    // sqrt(Ix^2 + Iy^2) > thresh with the hardware square root (1 ulp), which
    // that margin keeps below thresh.  Such a tile writes zero counts and
    // leaves; the barrier that ends the passes takes the vote.
    const float quiet = J.thresh * J.thresh * (1.0f - 0x1p-20f);
    if (!__syncthreads_or(!(J.thresh > 0.f) || mag2 > quiet)) {
        for (int r = threadIdx.x; r < 2 * TH; r += NT) {
            const int R = 2 * tr * TH + r;
            if (R < 2 * J.H) counts[J.counts_base + (long long)R * J.ntc + tc] = 0;
        }
        if (threadIdx.x < 4) tile_cnt[(long long)tile * 4 + threadIdx.x] = 0;
        return;
    }
    // the vertical-pass planes are dead now: the survivor queue lives in their place
    int *queue = (int *)(lds + Gm::OFF_P0);
    nms_tile<R1, TH, EB>(lds, J, tr, tc, tile_local, counts, stg, sh_cnt, (int *)(lds + Gm::OFF_CAND), queue, rowmask);
    // each survivor's raster key (local row, local column, rank in its row):
    // the row masks are final after the barrier inside nms_tile
    {
        const int qn = sh_cnt[4];
        for (int e = threadIdx.x; e < qn; e += NT) {
            const int q = queue[e];
            const int r = q & 63, c = (q >> 6) & 63, slot = q >> 12;
            const int rank = __popcll(rowmask[r] & ((1ull << c) - 1ull));
            stg.key[J.stage_base + tile_local * 4 * J.stage_cap + slot] =
                (uint32_t)r | ((uint32_t)c << 8) | ((uint32_t)rank << 16);
        }
    }
    if (threadIdx.x < 4) {
        const int n = sh_cnt[threadIdx.x];
        tile_cnt[(long long)tile * 4 + threadIdx.x] = n;
        if (n > J.stage_cap && J.flags_out) atomicOr(J.flags_out, (int)TOED_FLAG_STAGING);
    }
    // a tile with edgels joins the list the finalize kernel walks
    // (tlist[0] entries from tlist[1], in no particular order)
    if (threadIdx.x == 0 && (sh_cnt[0] | sh_cnt[1] | sh_cnt[2] | sh_cnt[3])) tlist[1 + atomicAdd(tlist, 1)] = tile;
}

// Taps of one sigma passed by value (the forward detection's part of a
// TapSet), so that the allocation-free entry point can place them in its
// static scratch array with a kernel launch instead of a copy.
struct TapArgs {
    int R0, R1;
    float fu[2][MAXTAPS];
    float fs[2][MAXTAPS];
    double du[4][MAXTAPS];
    double ds[4][MAXTAPS];
};

__global__ __launch_bounds__(256) void toed_put_taps_kernel(TapSet *dst, TapArgs a) {
    const int t = threadIdx.x;
    if (t == 0) { dst->R0 = a.R0; dst->R1 = a.R1; dst->pad0 = 0; dst->pad1 = 0; }
    if (t < 2 * MAXTAPS) {
        const int f = t / MAXTAPS, k = t % MAXTAPS;
        dst->fu[f][k] = a.fu[f][k];
        dst->fs[f][k] = a.fs[f][k];
    }
    if (t < 4 * MAXTAPS) {
        const int f = t / MAXTAPS, k = t % MAXTAPS;
        dst->du[f][k] = a.du[f][k];
        dst->ds[f][k] = a.ds[f][k];
        dst->dsu[f][k] = 0.0;    // the sigma derivatives belong to the backward, which this form has not
        dst->dss[f][k] = 0.0;
    }
}

// The reference launch layer's eight device tap arrays (gpu_kernels.hpp:24-45:
// G', G'', G''', G unshifted, then shifted, 19 taps each for its fixed 17-tap
// window, gpu_toed.hpp:368-386) packed into a TapSet (R0 = 8, R1 = 9).
template <typename T>
__global__ __launch_bounds__(64) void toed_pack_taps_kernel(TapSet *dst, const T *Gx, const T *Gxx, const T *Gxxx,
                                                            const T *G, const T *Gx_sh, const T *Gxx_sh,
                                                            const T *Gxxx_sh, const T *G_sh) {
    const int k = threadIdx.x;
    if (k == 0) { dst->R0 = MAXR1 - 1; dst->R1 = MAXR1; dst->pad0 = 0; dst->pad1 = 0; }
    if (k < MAXTAPS) {
        dst->du[0][k] = (double)G[k];    dst->du[1][k] = (double)Gx[k];
        dst->du[2][k] = (double)Gxx[k];  dst->du[3][k] = (double)Gxxx[k];
        dst->ds[0][k] = (double)G_sh[k]; dst->ds[1][k] = (double)Gx_sh[k];
        dst->ds[2][k] = (double)Gxx_sh[k]; dst->ds[3][k] = (double)Gxxx_sh[k];
        dst->fu[0][k] = (float)G[k];     dst->fu[1][k] = (float)Gx[k];
        dst->fs[0][k] = (float)G_sh[k];  dst->fs[1][k] = (float)Gx_sh[k];
        for (int f = 0; f < 4; ++f) { dst->dsu[f][k] = 0.0; dst->dss[f][k] = 0.0; }
    }
}

// K0: per-image flag reset and the count of the list of tiles with edgels
// (a kernel, not a memset: safe inside captured graphs)
__global__ void toed_reset_kernel(const ImageJob *jobs, int njobs, ImageJob job0, int *tlist) {
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m == 0) *tlist = 0;
    if (m >= njobs) return;
    const ImageJob &J = jobs ? jobs[m] : job0;
    if (J.flags_out) *J.flags_out = 0;
}

// K1: one workgroup per tile (plain form) or a persistent grid pulling tiles
// from a device counter (re-armed by K3's last workgroup).  One
// instantiation per window radius R1; a launch covers the tiles
// [tile_begin, tile_end) of the images with that radius.
// Five workgroups per CU for u8 images (byte-staged image) and for 8-row
// tiles; four for 16-row tiles of f32 images at sigma 2 (LDS).
template <typename IN_T, int TH, int R1>
__global__ __launch_bounds__(NT, (TH == 8 || stage_bytes<IN_T>() == 1) ? 5 : 4) void toed_tiles_kernel(
    const ImageJob *jobs, int njobs, ImageJob job0, int *counts, StageBuf staging, int *tile_cnt, int *tlist,
    int tile_begin, int tile_end, int *ticket) {
    constexpr int EB = stage_bytes<IN_T>();
    __shared__ float lds[Geo<R1, TH, EB>::TOTAL];
    __shared__ int sh_cnt[6];
    __shared__ int sh_tile;
    if (ticket == nullptr) {
        const int tile = tile_begin + blockIdx.x;
        const ImageJob J = load_job(jobs, njobs, job0, tile);
        tile_setup<R1, TH, EB>(lds, sh_cnt);
        stage_tile<IN_T, R1, TH>(lds, J, tile);
        __syncthreads();
        tile_compute<IN_T, R1, TH>(lds, sh_cnt, J, tile, counts, staging, tile_cnt, tlist);
        return;
    }
    for (;;) {
        if (threadIdx.x == 0) sh_tile = tile_begin + atomicAdd(ticket, 1);
        __syncthreads();
        const int tile = sh_tile;
        __syncthreads();
        if (tile >= tile_end) break;
        const ImageJob J = load_job(jobs, njobs, job0, tile);
        tile_setup<R1, TH, EB>(lds, sh_cnt);
        stage_tile<IN_T, R1, TH>(lds, J, tile);
        __syncthreads();
        tile_compute<IN_T, R1, TH>(lds, sh_cnt, J, tile, counts, staging, tile_cnt, tlist);
        __syncthreads();
    }
}

// K2a: exclusive scan of one CHUNK of counts; chunk totals to chunk_sum.
__global__ __launch_bounds__(256) void toed_scan_chunks(const ImageJob *jobs, int njobs, ImageJob job0,
                                                        const int *counts, int *offsets, int *chunk_sum) {
    __shared__ int wsum[4];
    const long long chunk = blockIdx.x;
    const int m = jobs ? job_for_chunk(jobs, njobs, chunk) : 0;
    const ImageJob &J = jobs ? jobs[m] : job0;
    const long long nvalid = (long long)(2 * J.H) * J.ntc;
    const long long local0 = (chunk - J.chunk_base) * CHUNK;   // entry index within the image
    constexpr int PER = CHUNK / 256;
    int v[PER];
    int s = 0;
    const long long e0 = local0 + threadIdx.x * PER;
#pragma unroll
    for (int k = 0; k < PER; ++k) {
        const long long e = e0 + k;
        v[k] = (e < nvalid) ? counts[J.counts_base + e] : 0;
        s += v[k];
    }
    // wavefront inclusive scan of per-thread sums
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int inc = s;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const int y = __shfl_up(inc, d);
        if (lane >= d) inc += y;
    }
    if (lane == 63) wsum[wave] = inc;
    __syncthreads();
    int wbase = 0;
    for (int w = 0; w < wave; ++w) wbase += wsum[w];
    int run = wbase + inc - s;
#pragma unroll
    for (int k = 0; k < PER; ++k) {
        const long long e = e0 + k;
        if (e < nvalid) offsets[J.counts_base + e] = run;
        run += v[k];
    }
    if (threadIdx.x == 255) chunk_sum[chunk] = wbase + inc;
}

// K2b: one workgroup per image: exclusive scan of its chunk totals; N out.
__global__ __launch_bounds__(1024) void toed_scan_images(const ImageJob *jobs, int njobs, ImageJob job0,
                                                         int *chunk_sum) {
    __shared__ int wsum[16];
    __shared__ int carry;
    const ImageJob &J = jobs ? jobs[blockIdx.x] : job0;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int base = 0; base < J.nchunks; base += 1024) {
        const int k = base + threadIdx.x;
        const int x = (k < J.nchunks) ? chunk_sum[J.chunk_base + k] : 0;
        int inc = x;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const int y = __shfl_up(inc, d);
            if (lane >= d) inc += y;
        }
        if (lane == 63) wsum[wave] = inc;
        __syncthreads();
        int wbase = carry;
        for (int w = 0; w < wave; ++w) wbase += wsum[w];
        if (k < J.nchunks) chunk_sum[J.chunk_base + k] = wbase + inc - x;
        __syncthreads();
        if (threadIdx.x == 1023) carry = wbase + inc;
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const int n = carry;
        if (J.count_out) *J.count_out = n;
        if (!J.cbase && n > J.capacity && J.flags_out) atomicOr(J.flags_out, (int)TOED_FLAG_CAPACITY);
    }
}

// Orientation mode 1: the nine responses are evaluated at the edgel's
// subpixel location instead of at its interp pixel, by cubic convolution
// (Keys, a = -1/2) of the interp-grid responses of the 4 x 4 interp pixels
// around it -- the MATLAB detector's QaD path (third_order_edge_detector.m:
// 91-121, interp2 'cubic'), then the same third-order formula
// (cpu_toed.cpp:227-232).  Each of the 16 responses is the FP64 separable
// sum of the phase's taps over the image window (zero outside the image), as
// the orientation pass computes it at the pixel.
__device__ __forceinline__ double keys_w(double s) {
    const double x = fabs(s);
    if (x <= 1.0) return (1.5 * x - 2.5) * x * x + 1.0;
    if (x < 2.0) return ((-0.5 * x + 2.5) * x - 4.0) * x + 2.0;
    return 0.0;
}

// K2b (orientation mode 1 only): per tile, after the tile kernel and before
// the scatter, each staged edgel's orientation is recomputed at its subpixel
// location, one wavefront per edgel.  With Xs = 2x + 1, Ys = 2y + 1 the
// stencil's interp pixels are (Rp, Cp) = (floor(Ys) - 1 + i, floor(Xs) - 1 + j),
// i, j = 0..3, on image pixel (Rp >> 1, Cp >> 1) with phase (Rp & 1, Cp & 1).
// The 16 responses share one union window of at most 2 R1 + 3 image rows, so
// the column sums are formed once per (window row, stencil column):
//   (1) lane (h, r) takes window row r and stencil columns 2h and 2h + 1 (one
//       even and one odd interp column; for the even one also the sum without
//       the two end taps, phase (0, 0)'s R0 window) and leaves them in LDS;
//   (2) lanes 0..15, one per stencil point, sum their column's sums against
//       the row taps and apply the point's Keys weight;
//   (3) the 16 weighted responses are summed with DPP.
// Each sum runs over the taps from -R1 to R1 in the same order as a direct
// per-point evaluation would.
constexpr int QROWS = 2 * MAXR1 + 3;
constexpr int QWAVES = NT / 64;

__global__ __launch_bounds__(NT) void toed_qad_kernel(const ImageJob *jobs, int njobs, ImageJob job0, StageBuf stg,
                                                      const int *tile_cnt) {
    // per wavefront: full sums [stencil column][filter][window row], and the
    // R0-window sums of the even column of each pair [pair][filter][window row]
    __shared__ double hs[QWAVES][4][4][QROWS];
    __shared__ double hr[QWAVES][2][4][QROWS];
    __shared__ double ytaps[2][4][MAXTAPS];     // du, ds: the row taps of the vertical pass
    const int tile = blockIdx.x;
    const ImageJob J = load_job(jobs, njobs, job0, tile);
    if (!J.qad) return;
    const TOED_CONST TapSet *tp = as_const(J.taps);
    for (int k = threadIdx.x; k < 8 * MAXTAPS; k += NT) {
        const int set = k / (4 * MAXTAPS), f = (k / MAXTAPS) & 3, t = k % MAXTAPS;
        ytaps[set][f][t] = set ? tp->ds[f][t] : tp->du[f][t];
    }
    __syncthreads();
    const int R1 = J.R1;
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    double(*Hs)[4][QROWS] = hs[wave];
    double(*Hr)[4][QROWS] = hr[wave];
    const long long tile_local = tile - J.tile_begin;
    // the tile's edgels of all four phases are dealt to the wavefronts in turn
    const int n0 = min(tile_cnt[(long long)tile * 4 + 0], J.stage_cap);
    const int n1 = min(tile_cnt[(long long)tile * 4 + 1], J.stage_cap);
    const int n2 = min(tile_cnt[(long long)tile * 4 + 2], J.stage_cap);
    const int n3 = min(tile_cnt[(long long)tile * 4 + 3], J.stage_cap);
    const int total = n0 + n1 + n2 + n3;
    for (int k = wave; k < total; k += QWAVES) {
        int phase = 0, e = k;
        if (e >= n0) {
            e -= n0;
            phase = 1;
            if (e >= n1) {
                e -= n1;
                phase = 2;
                if (e >= n2) {
                    e -= n2;
                    phase = 3;
                }
            }
        }
        const long long base = J.stage_base + (tile_local * 4 + phase) * J.stage_cap;
        const float4 rec = stg.rec[base + e];
        const double Xs = 2.0 * (double)rec.x + 1.0, Ys = 2.0 * (double)rec.y + 1.0;
        const int rx = __builtin_amdgcn_readfirstlane((int)floor(Xs));
        const int ry = __builtin_amdgcn_readfirstlane((int)floor(Ys));
        const int ci0 = (ry - 1) >> 1;      // image row of stencil row 0
        {
            const int r = lane & 31, hh = lane >> 5;
            // d = 1 when stencil column 0 is odd: the even column of each
            // pair is then column 2h + 1, one image pixel right of the odd one
            const int d = (rx - 1) & 1;
            const int je = 2 * hh + d, jo = 2 * hh + 1 - d;
            const int co = (rx - 1 + jo) >> 1;
            const int gi = ci0 - R1 + r;
            double fe[4] = {0, 0, 0, 0}, ie[4] = {0, 0, 0, 0}, fo[4] = {0, 0, 0, 0};
            if (r < QROWS && gi >= 0 && gi < J.H) {
                const long long row = (long long)gi * J.pitch;
                // image column co + R1 + d - m: tap m - R1 of the even
                // column (m <= 2 R1), tap m - d - R1 of the odd one (m >= d)
                for (int m = 0; m <= 2 * R1 + d; ++m) {
                    const int gj = co + R1 + d - m;
                    double v = 0.0;
                    if (gj >= 0 && gj < J.W)
                        v = J.in_type == TOED_U8 ? (double)((const uint8_t *)J.img)[row + gj]
                                                 : (double)((const float *)J.img)[row + gj];
                    if (m <= 2 * R1) {
#pragma unroll
                        for (int f = 0; f < 4; ++f) fe[f] = fma(v, tp->du[f][m], fe[f]);
                        if (m != 0 && m != 2 * R1) {
#pragma unroll
                            for (int f = 0; f < 4; ++f) ie[f] = fma(v, tp->du[f][m], ie[f]);
                        }
                    }
                    if (m >= d) {
#pragma unroll
                        for (int f = 0; f < 4; ++f) fo[f] = fma(v, tp->ds[f][m - d], fo[f]);
                    }
                }
            }
            if (r < QROWS) {
#pragma unroll
                for (int f = 0; f < 4; ++f) {
                    Hs[je][f][r] = fe[f];
                    Hr[hh][f][r] = ie[f];
                    Hs[jo][f][r] = fo[f];
                }
            }
        }
        // the sums written above are read by other lanes of this wavefront
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        double r9[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        if (lane < 16) {
            const int i = lane >> 2, j = lane & 3;
            const double w = keys_w(Ys - ry + 1 - i) * keys_w(Xs - rx + 1 - j);
            const int Rp = ry - 1 + i, Cp = rx - 1 + j, a = Rp & 1, b = Cp & 1, ci = Rp >> 1;
            const int lo = (a == 0 && b == 0) ? 1 : 0;
            const double(*Y)[MAXTAPS] = ytaps[a];
            const double(*Hj)[QROWS] = lo ? Hr[j >> 1] : Hs[j];
            if (w != 0.0) {
                for (int p = -R1 + lo; p <= R1 - lo; ++p) {
                    const int gi = ci - p;
                    if (gi < 0 || gi >= J.H) continue;
                    const int rr = gi - (ci0 - R1);
                    const double h0 = Hj[0][rr], h1 = Hj[1][rr], h2 = Hj[2][rr], h3 = Hj[3][rr];
                    const int t = p + R1;
                    const double y0 = Y[0][t], y1 = Y[1][t], y2 = Y[2][t], y3 = Y[3][t];
                    r9[0] = fma(h1, y0, r9[0]);   // fx
                    r9[1] = fma(h0, y1, r9[1]);   // fy
                    r9[2] = fma(h2, y0, r9[2]);   // fxx
                    r9[3] = fma(h1, y1, r9[3]);   // fxy
                    r9[4] = fma(h0, y2, r9[4]);   // fyy
                    r9[5] = fma(h2, y1, r9[5]);   // fxxy
                    r9[6] = fma(h1, y2, r9[6]);   // fxyy
                    r9[7] = fma(h3, y0, r9[7]);   // fxxx
                    r9[8] = fma(h0, y3, r9[8]);   // fyyy
                }
            }
#pragma unroll
            for (int k = 0; k < 9; ++k) r9[k] *= w;
        }
#pragma unroll
        for (int k = 0; k < 9; ++k) r9[k] = toeddpp::group_sum<16>(r9[k]);
        if (lane == 0) {
            const double fx = r9[0], fy = r9[1], fxx = r9[2], fxy = r9[3], fyy = r9[4];
            const double fxxy = r9[5], fxyy = r9[6], fxxx = r9[7], fyyy = r9[8];
            double tx = fx * (2 * fxx * fxx + 2 * fxy * fxy) + fy * (2 * fxx * fxy + 2 * fyy * fxy)
                      + 2 * fx * fy * fxxy + fy * fy * fxyy + fx * fx * fxxx;
            double ty = fx * (2 * fxx * fxy + 2 * fyy * fxy) + fy * (2 * fyy * fyy + 2 * fxy * fxy)
                      + 2 * fx * fy * fxyy + fx * fx * fxxy + fy * fy * fyyy;
            stg.rec[base + e].z = to_angle(tx, ty);
        }
        // the next edgel's sums overwrite this one's only after they were read
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
}

// K3: the third-order orientation of every staged edgel (theta_group, 8 lanes
// per edgel; in orientation mode 1 toed_qad_kernel has already set it), then
// the edgel to offset[row, tile column] + rank, the reference's raster order of
// the interp grid (cpu_toed.cpp:510-538).  One wavefront per tile (a tile
// stages a handful of edgels per phase).
__device__ __forceinline__ long long out_index(const ImageJob &J, const int *offsets, const int *chunk_pref, int R,
                                               int tc, int rank) {
    const long long ce = J.counts_base + (long long)R * J.ntc + tc;
    return (long long)offsets[ce] + chunk_pref[J.chunk_base + (ce - J.counts_base) / CHUNK] + rank;
}

// Output row of an edgel: its rank in the image's list, plus, with compact
// output, the image's first row in the shared buffer.
__device__ __forceinline__ long long out_row(const ImageJob &J, long long idx) {
    return J.cbase ? idx + *J.cbase : idx;
}

// Compact output: image m's rows start after those of images 0 .. m-1 (the
// caller's order).  One workgroup scans the per-image counts; an image whose
// rows would pass the shared capacity gets TOED_FLAG_CAPACITY.
__global__ __launch_bounds__(1024) void toed_compact_base_kernel(const int *counts, int *flags, int n,
                                                                 long long capacity, long long *cbase) {
    __shared__ long long wsum[16];
    __shared__ long long carry;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int base = 0; base < n; base += 1024) {
        const int m = base + threadIdx.x;
        const long long x = m < n ? counts[m] : 0;
        long long inc = x;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const long long y = __shfl_up(inc, d);
            if (lane >= d) inc += y;
        }
        if (lane == 63) wsum[wave] = inc;
        __syncthreads();
        long long wb = carry;
        for (int w = 0; w < wave; ++w) wb += wsum[w];
        if (m < n) {
            cbase[m] = wb + inc - x;
            if (wb + inc > capacity && flags) atomicOr(flags + m, (int)TOED_FLAG_CAPACITY);
        }
        __syncthreads();
        if (threadIdx.x == 1023) carry = wb + inc;
        __syncthreads();
    }
}

// A grid of at most FIN_GRID workgroups of four wavefronts walks the list of
// tiles with edgels that the tile kernel built (tiles without edgels cost
// nothing here); a tile's edgels are dealt to the workgroup's 32 groups of 8
// lanes, so the tiles with many edgels (along the scene's contours) finish in
// few rounds.
constexpr int FIN_NT = 256;
constexpr int FIN_GT = 8;                  // lanes per edgel
constexpr int FIN_G = FIN_NT / FIN_GT;     // lane groups per workgroup
constexpr int FIN_GRID = 2048;

__device__ __forceinline__ void finalize_tile(const ImageJob *jobs, int njobs, const ImageJob &job0, int TH,
                                              const int *offsets, const int *chunk_pref, StageBuf stg,
                                              const int *tile_cnt, int tile, double (*tapd)[MAXTAPS][4],
                                              double *fimg, float4 *s_rec, long long *s_idx, int *s_rc) {
    const ImageJob J = load_job(jobs, njobs, job0, tile);
    const long long tile_local = tile - J.tile_begin;
    const int tr = (int)(tile_local / J.ntc), tc = (int)(tile_local % J.ntc);
    const int IW = 2 * J.W;
    int nph[4];
#pragma unroll
    for (int ph = 0; ph < 4; ++ph) nph[ph] = min(tile_cnt[(long long)tile * 4 + ph], J.stage_cap);
    const int total = nph[0] + nph[1] + nph[2] + nph[3];
    if (total == 0) return;    // the same for the whole workgroup
    if (J.qad) {
        // orientation mode 1: toed_qad_kernel has set it; scatter only
        for (int e = threadIdx.x; e < total; e += FIN_NT) {
            int k = e, phase = 0;
            while (phase < 3 && k >= nph[phase]) { k -= nph[phase]; ++phase; }
            const long long gi = J.stage_base + (tile_local * 4 + phase) * J.stage_cap + k;
            const uint32_t key = stg.key[gi];
            const int r = key & 0xff, c = (key >> 8) & 0xff, rank = key >> 16;
            const int R = 2 * tr * TH + r, C = 2 * tc * TW + c;
            const long long idx = out_row(J, out_index(J, offsets, chunk_pref, R, tc, rank));
            if (idx >= J.capacity) continue;
            reinterpret_cast<float4 *>(J.out)[idx] = stg.rec[gi];
            if (J.index_out) J.index_out[idx] = R * IW + C;
        }
        return;
    }
    // staging slot of the tile's edgel e (phases in order)
    auto stage_slot = [&](int e) -> long long {
        int k = e, phase = 0;
        while (phase < 3 && k >= nph[phase]) { k -= nph[phase]; ++phase; }
        return J.stage_base + (tile_local * 4 + phase) * J.stage_cap + k;
    };
    // The first chunk's keys and records are loaded in the same round as the
    // FP64 taps of this tile's sigma and the tile's image region (rows and
    // columns from 1 + R1 before the tile, zero outside the image); only the
    // output rows, which need the keys, take a second round.
    const bool have0 = (int)threadIdx.x < total;
    const long long slot0 = stage_slot(have0 ? (int)threadIdx.x : 0);
    const uint32_t key0 = stg.key[slot0];
    const float4 rec0 = stg.rec[slot0];
    const TOED_CONST TapSet *tp = as_const(J.taps);
    const int R1 = J.R1, N = 2 * R1 + 1;
    for (int k = threadIdx.x; k < 12 * N; k += FIN_NT) {
        const int fam = k / (4 * N), f = (k / N) & 3, t = k % N;
        double v = (fam == 2) ? tp->ds[f][t] : tp->du[f][t];
        if (fam == 1 && (t == 0 || t == N - 1)) v = 0.0;
        tapd[fam][t][f] = v;
    }
    {
        const int nri = TH + 2 + 2 * R1, nci = TW + 2 + 2 * R1;
        const int gi0 = tr * TH - 1 - R1, gj0 = tc * TW - 1 - R1;
        constexpr int RPP = FIN_NT / 64;                              // rows per pass, a lane per column
        const int col = threadIdx.x & 63, r0 = threadIdx.x >> 6;
        const int gj = gj0 + col;
        const bool colok = col < nci && gj >= 0 && gj < J.W;
        const int gjc = min(max(gj, 0), J.W - 1);
        constexpr int NPASS = (FROWS + RPP - 1) / RPP;
        float vals[NPASS];      // every row's load issued before the first store
        if (J.in_type == TOED_U8) {
            const uint8_t *src = (const uint8_t *)J.img + gjc;
#pragma unroll
            for (int u = 0; u < NPASS; ++u)
                vals[u] = (float)src[(long long)min(max(gi0 + r0 + RPP * u, 0), J.H - 1) * J.pitch];
        } else {
            const float *src = (const float *)J.img + gjc;
#pragma unroll
            for (int u = 0; u < NPASS; ++u) vals[u] = src[(long long)min(max(gi0 + r0 + RPP * u, 0), J.H - 1) * J.pitch];
        }
        if (col < nci) {
#pragma unroll
            for (int u = 0; u < NPASS; ++u) {
                const int rr = r0 + RPP * u, gi = gi0 + rr;
                if (rr < nri) fimg[rr * FLDI + col] = (colok && gi >= 0 && gi < J.H) ? (double)vals[u] : 0.0;
            }
        }
    }
    // a chunk of staged edgels to LDS: record, output row and tile-local pixel
    auto put_edgel = [&](uint32_t key, float4 rec) {
        const int r = key & 0xff, c = (key >> 8) & 0xff, rank = key >> 16;
        const long long idx = out_row(J, out_index(J, offsets, chunk_pref, 2 * tr * TH + r, tc, rank));
        s_rec[threadIdx.x] = rec;
        s_idx[threadIdx.x] = idx < J.capacity ? idx : -1;
        s_rc[threadIdx.x] = r | (c << 8);
    };
    if (have0) put_edgel(key0, rec0);
    const int g = threadIdx.x / FIN_GT, l = threadIdx.x % FIN_GT;
    for (int c0 = 0; c0 < total; c0 += FIN_NT) {
        if (c0 > 0) {
            const int e = c0 + threadIdx.x;
            if (e < total) {
                const long long gi = stage_slot(e);
                put_edgel(stg.key[gi], stg.rec[gi]);
            }
        }
        __syncthreads();
        const int cn = min(FIN_NT, total - c0);
        for (int e0 = 0; e0 < cn; e0 += FIN_G) {
            const int e = e0 + g;
            if (e >= cn) break;                 // the same on the lanes of a group
            const int rc = s_rc[e];
            const int r = rc & 0xff, c = rc >> 8;
            const int ci = (r >> 1) + 1 + R1, cj = (c >> 1) + 1 + R1, a = r & 1, b = c & 1;
            double th;
            if (R1 == 5) th = theta_group<5, FIN_GT>(fimg, ci, cj, a, b, l, tapd);
            else if (R1 == 7) th = theta_group<7, FIN_GT>(fimg, ci, cj, a, b, l, tapd);
            else th = theta_group<9, FIN_GT>(fimg, ci, cj, a, b, l, tapd);
            if (l == 0) {
                const long long idx = s_idx[e];
                if (idx >= 0) {
                    float4 rec = s_rec[e];
                    rec.z = (float)th;
                    reinterpret_cast<float4 *>(J.out)[idx] = rec;
                    if (J.index_out) J.index_out[idx] = (2 * tr * TH + r) * IW + 2 * tc * TW + c;
                }
            }
        }
        __syncthreads();   // the next chunk (or tile) overwrites this chunk's slots only after they were read
    }
}

__global__ __launch_bounds__(FIN_NT) void toed_finalize_kernel(const ImageJob *jobs, int njobs, ImageJob job0, int TH,
                                                               const int *offsets, const int *chunk_pref,
                                                               StageBuf stg, const int *tile_cnt, const int *tlist,
                                                               int *ticket) {
    __shared__ double tapd[3][MAXTAPS][4];
    __shared__ double fimg[FROWS * FLDI];
    __shared__ float4 s_rec[FIN_NT];       // a chunk of staged edgels: record,
    __shared__ long long s_idx[FIN_NT];    // output row (-1: none),
    __shared__ int s_rc[FIN_NT];           // tile-local interp row | col << 8
    // the persistent tile kernels' tickets are re-armed once every tile is final
    if (ticket && blockIdx.x == gridDim.x - 1 && threadIdx.x < NGROUPS) ticket[threadIdx.x] = 0;
    const int nlist = tlist[0];
    for (int k = blockIdx.x; k < nlist; k += gridDim.x)
        finalize_tile(jobs, njobs, job0, TH, offsets, chunk_pref, stg, tile_cnt, tlist[1 + k], tapd, fimg, s_rec,
                      s_idx, s_rc);
}

// ---------------------------------------------------------------------------
// Backward of a batch plan's last forward, tile by tile (differentiable form)
// ---------------------------------------------------------------------------
// One workgroup per forward tile: the tile's image plus a (R1+1)-pixel halo is
// staged in LDS once, every survivor of the tile (from the plan's staging, or
// from its interp rows when the plan keeps the interp index) runs the shared
// FP64 reverse pass on a lane group, its transposed filters are added by one
// wavefront into that wavefront's own (TH + 2R1 + 2) x (32 + 2R1 + 2)
// gradient patch, and the summed patches are added to dL/dimage once per
// tile.  dL/dsigma is summed per workgroup.
// Occupancy 4 waves per SIMD (128 VGPRs, a few spills to scratch) measures
// faster than the 2 that the unconstrained allocation (about 230 VGPRs) gives.
template <typename IN_T, int TH, int R1>
__global__ __launch_bounds__(NT, 4) void toed_tile_backward_kernel(const ImageJob *jobs, int njobs, ImageJob job0,
                                                                const GradOut *gouts, StageBuf stg,
                                                                const int *tile_cnt, const int *counts,
                                                                const int *offsets, const int *chunk_pref,
                                                                int tile_begin) {
    constexpr int N = 2 * R1 + 1;
    constexpr int U = R1 + 1;
    constexpr int PR = TH + 2 * U, PC = TW + 2 * U, LDG = PC | 1;
    constexpr int NW = NT / 64;              // wavefronts
    __shared__ float simg[PR * LDG];
    __shared__ float sgrad[NW][PR * LDG];    // one gradient patch per wavefront
    __shared__ double stp[3][N + 2][4];
    __shared__ double sdp[3][N + 2][4];
    __shared__ float4 stf4[3][N + 2];
    __shared__ double sgs[NT / BW_G];
    __shared__ int s_rc[BW_CHUNK];           // tile-local interp (row | col << 8) of a chunk's edgels, -1: no output row
    __shared__ float4 s_up[BW_CHUNK];        // their upstream gradient rows
    constexpr int NG = NT / BW_G;            // edgels per round (one per lane group)
    __shared__ float s_adj[NG][toedadj::NADJ];
    __shared__ int s_geo[NG];                // tile-local (row | col << 8 | octant << 16) of a round's edgels, -1: none
    __shared__ int s_rowpre[2 * TH + 1];     // edgels of the tile's interp rows before each row
    __shared__ long long s_rowstart[2 * TH]; // output row (image-relative) of each interp row's first edgel
    const int tile = tile_begin + blockIdx.x;
    const int m = job_index(jobs, njobs, tile);
    const ImageJob J = load_job(jobs, njobs, job0, tile);
    const GradOut O = gouts[m];
    const long long tile_local = tile - J.tile_begin;
    const int tr = (int)(tile_local / J.ntc), tc = (int)(tile_local % J.ntc);
    const int i0 = tr * TH, j0 = tc * TW;
    const int gi0 = i0 - U, gj0 = j0 - U;        // patch origin (image rows/cols)
    const TOED_CONST TapSet *tp = as_const(J.taps);
    // Prologue: when the plan keeps each edgel's interp index, the count and
    // output start of each of the tile's interp rows (an empty tile returns
    // here); then the tile's independent global loads in one pass so that
    // their latencies overlap: the tap tables (threads < 3 (2 R1 + 3); variant
    // 1, window R0, is variant 0 with zero end taps) and the image region.
    constexpr int NP = N + 2;
    const bool by_rows = J.index_out != nullptr;
    int rcnt = 0;
    if (by_rows && threadIdx.x < 2 * TH && 2 * i0 + (int)threadIdx.x < 2 * J.H) {   // rows of the image only
        const long long ce = J.counts_base + (long long)(2 * i0 + threadIdx.x) * J.ntc + tc;
        rcnt = counts[ce];
        s_rowstart[threadIdx.x] = (long long)offsets[ce] + chunk_pref[J.chunk_base + (ce - J.counts_base) / CHUNK];
    }
    // a tile without edgels adds nothing: it leaves before staging anything
    {
        bool any;
        if (by_rows) {
            any = __syncthreads_or(rcnt > 0) != 0;
        } else {
            int t = 0;
#pragma unroll
            for (int ph = 0; ph < 4; ++ph) t += tile_cnt[(long long)tile * 4 + ph];
            any = t > 0;
        }
        if (!any) return;
    }
    double ta[4], td[4];
    const int tv = threadIdx.x / NP, ti = threadIdx.x - tv * NP, tt = ti - 1;
    if (threadIdx.x < 3 * NP) {
        const bool in = tt >= 0 && tt <= 2 * R1 && !(tv == 1 && (tt == 0 || tt == N - 1));
#pragma unroll
        for (int f = 0; f < 4; ++f) {
            ta[f] = in ? (tv == 2 ? tp->ds[f][tt] : tp->du[f][tt]) : 0.0;
            td[f] = in ? (tv == 2 ? tp->dss[f][tt] : tp->dsu[f][tt]) : 0.0;
        }
    }
    for (int e = threadIdx.x; e < PR * PC; e += NT) {
        const int r = e / PC, c = e % PC;
        const int gi = gi0 + r, gj = gj0 + c;
        float v = 0.f;
        if (gi >= 0 && gi < J.H && gj >= 0 && gj < J.W) v = load_px<IN_T>(J.img, (long long)gi * J.pitch + gj);
        simg[r * LDG + c] = v;
#pragma unroll
        for (int w = 0; w < NW; ++w) sgrad[w][r * LDG + c] = 0.f;
    }
    if (threadIdx.x < 3 * NP) {
#pragma unroll
        for (int f = 0; f < 4; ++f) {
            stp[tv][ti][f] = ta[f];
            sdp[tv][ti][f] = td[f];
        }
        stf4[tv][ti] = make_float4((float)ta[0], (float)ta[1], (float)ta[2], (float)ta[3]);
    }
    if (by_rows && threadIdx.x < 2 * TH) s_rowpre[threadIdx.x + 1] = rcnt;
    if (threadIdx.x == 0) s_rowpre[0] = 0;
    const int g = threadIdx.x / BW_G, l = threadIdx.x % BW_G;
    if (l == 0) sgs[g] = 0.0;
    __syncthreads();
    struct LdsImg {
        const float *p; int gi0, gj0;
        __device__ double operator()(int i, int j) const { return (double)p[(i - gi0) * LDG + (j - gj0)]; }
    } img{simg, gi0, gj0};
    // All four phases form one work list, processed in chunks: a cooperative
    // pre-pass resolves each edgel's interp pixel, output row and upstream
    // gradient into LDS (no global-load chains inside the per-edgel work).
    // Each chunk goes in rounds of NG edgels, one per lane group, spread over
    // the wavefronts.  A group runs the FP64 reverse pass of its edgel and
    // leaves the adjoint coefficients in LDS; then its wavefront scatters its
    // groups' edgels one after the other into its own gradient patch, each
    // edgel by all 64 lanes (a lane takes a row and four columns of the
    // edgel's patch), with plain read-add-writes: the lanes of one edgel
    // touch distinct pixels, a wavefront's LDS operations complete in order,
    // and no other wavefront writes that patch (LDS float atomics cost
    // several cycles per lane).  The patches are summed at the end.
    int nph[4];
#pragma unroll
    for (int ph = 0; ph < 4; ++ph) nph[ph] = min(tile_cnt[(long long)tile * 4 + ph], J.stage_cap);
    int total = nph[0] + nph[1] + nph[2] + nph[3];
    if (by_rows) {
        // inclusive prefix of the row counts: each row's thread sums the rows before it
        int pre = 0;
        if (threadIdx.x <= 2 * TH)
            for (int u = 1; u <= threadIdx.x; ++u) pre += s_rowpre[u];
        __syncthreads();
        if (threadIdx.x <= 2 * TH) s_rowpre[threadIdx.x] = pre;
        __syncthreads();
        total = s_rowpre[2 * TH];
    }
    constexpr int GPW = 64 / BW_G;                 // groups per wavefront
    constexpr int P = 2 * U;                       // patch side of one edgel
    constexpr int PQ = (P + 3) / 4;                // four-column blocks per patch row
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int slot = (g % GPW) * NW + wave;        // this group's edgel within a round
    float *wgrad = sgrad[wave];
    for (int c0 = 0; c0 < total; c0 += BW_CHUNK) {
        const int cn = min(BW_CHUNK, total - c0);
        for (int k = threadIdx.x; k < cn; k += NT) {
            const int e = c0 + k;
            long long idx;
            int r = 0, c = 0;
            if (by_rows) {
                // the interp row holding edgel e of the tile, then its output row
                int lo = 0, hi = 2 * TH - 1;
                while (lo < hi) {
                    const int mid = (lo + hi + 1) >> 1;
                    if (s_rowpre[mid] <= e) lo = mid; else hi = mid - 1;
                }
                r = lo;
                idx = s_rowstart[lo] + (e - s_rowpre[lo]);
                if (idx < J.capacity) c = J.index_out[out_row(J, idx)] % (2 * J.W) - 2 * j0;
            } else {
                int ee = e, phase = 0;
                while (ee >= nph[phase]) { ee -= nph[phase]; ++phase; }
                const uint32_t key = stg.key[J.stage_base + (tile_local * 4 + phase) * J.stage_cap + ee];
                r = key & 0xff;
                c = (key >> 8) & 0xff;
                idx = out_index(J, offsets, chunk_pref, 2 * i0 + r, tc, key >> 16);
            }
            s_rc[k] = idx < J.capacity ? (r | (c << 8)) : -1;
            s_up[k] = idx < J.capacity ? reinterpret_cast<const float4 *>(O.gout)[idx] : make_float4(0, 0, 0, 0);
        }
        __syncthreads();
        for (int r0 = 0; r0 < cn; r0 += NG) {
            const int nr = min(NG, cn - r0);
            if (slot < nr) {
                const int k = r0 + slot;
                int geo = -1;
                if (s_rc[k] >= 0) {
                    const int R = 2 * i0 + (s_rc[k] & 0xff), C = 2 * j0 + (s_rc[k] >> 8);
                    double gs = 0.0;
                    int oct;
                    // every lane of the group writes the same coefficients
                    if (toedadj::edgel_adjoint<R1, BW_G>(R, C, s_up[k], l, stp, sdp, img, s_adj[slot], &oct, &gs))
                        geo = s_rc[k] | (oct << 16);
                    if (l == 0) sgs[g] += gs;
                }
                if (l == 0) s_geo[slot] = geo;
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            // this wavefront's edgels of the round: slots j * NW + wave
            for (int j = 0; j < GPW; ++j) {
                const int q = j * NW + wave;
                if (q >= nr) break;
                const int geo = s_geo[q];
                if (geo < 0) continue;
                const int R = 2 * i0 + (geo & 0xff), C = 2 * j0 + ((geo >> 8) & 0xff), oct = geo >> 16;
                const int pr0 = (R >> 1) - U + (R & 1), pc0 = (C >> 1) - U + (C & 1);
                int vx[5], qb[5], vy[5], yb[5];
                toedadj::point_taps<R1>(R, C, oct, vx, qb, vy, yb);
                for (int t = lane; t < P * PQ; t += 64) {
                    const int r = t / PQ, cb = (t - r * PQ) * 4;
                    float cf[5][4];
#pragma unroll
                    for (int m = 0; m < 5; ++m) {
                        const float4 y4 = stf4[vy[m]][yb[m] - r];
                        const float yv[4] = {y4.x, y4.y, y4.z, y4.w};
                        toedadj::point_row_coefs(m, s_adj[q], yv, cf[m]);
                    }
                    float *row = wgrad + (pr0 + r - gi0) * LDG + (pc0 - gj0);
#pragma unroll
                    for (int cc = 0; cc < 4; ++cc) {
                        const int c = cb + cc;
                        if (c >= P) break;
                        row[c] += toedadj::patch_column<R1>(cf, vx, qb, c, stf4);
                    }
                }
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            }
        }
        __syncthreads();     // the chunk's s_rc / s_up are rewritten next
    }
    __syncthreads();
    for (int e = threadIdx.x; e < PR * PC; e += NT) {
        const int r = e / PC, c = e % PC;
        const int gi = gi0 + r, gj = gj0 + c;
        float v = sgrad[0][r * LDG + c];
#pragma unroll
        for (int w = 1; w < NW; ++w) v += sgrad[w][r * LDG + c];
        if (v != 0.f && gi >= 0 && gi < J.H && gj >= 0 && gj < J.W)
            unsafeAtomicAdd(O.gimg + (long long)gi * O.gpitch + gj, v);
    }
    if (threadIdx.x == 0 && O.gsigma) {
        double t = 0.0;
        for (int k = 0; k < NT / BW_G; ++k) t += sgs[k];
        if (t != 0.0) unsafeAtomicAdd(O.gsigma, t);
    }
}

// ---------------------------------------------------------------------------
// Unfused reference launch layer: interp planes and NMS maps
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void toed_planes_kernel(int H, int W, const float *img, float *Ix, float *Iy,
                                                          float *mag, float *orient, const TapSet *tp) {
    const long long o = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long IW = 2LL * W;
    if (o >= 4LL * H * W) return;
    const int R = (int)(o / IW), C = (int)(o % IW);
    const int i = R >> 1, j = C >> 1, a = R & 1, b = C & 1;
    const int R1 = tp->R1, rad = (a | b) ? tp->R1 : tp->R0;
    const float *X0 = b ? tp->fs[0] : tp->fu[0];
    const float *X1 = b ? tp->fs[1] : tp->fu[1];
    const float *Y0 = a ? tp->fs[0] : tp->fu[0];
    const float *Y1 = a ? tp->fs[1] : tp->fu[1];
    const double *XD = b ? &tp->ds[0][0] : &tp->du[0][0];
    const double *YD = a ? &tp->ds[0][0] : &tp->du[0][0];
    float fx = 0.f, fy = 0.f;
    float g[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int p = -rad; p <= rad; ++p) {
        const int ii = i - p;
        if (ii < 0 || ii >= H) continue;
        float h0 = 0, h1 = 0, h2 = 0, h3 = 0;
        for (int q = -rad; q <= rad; ++q) {
            const int jj = j - q;
            if (jj < 0 || jj >= W) continue;
            const float v = img[(long long)ii * W + jj];
            h0 = fmaf(v, X0[q + R1], h0);
            h1 = fmaf(v, X1[q + R1], h1);
            h2 = fmaf(v, (float)XD[2 * MAXTAPS + q + R1], h2);
            h3 = fmaf(v, (float)XD[3 * MAXTAPS + q + R1], h3);
        }
        fx = fmaf(h1, Y0[p + R1], fx);
        fy = fmaf(h0, Y1[p + R1], fy);
        const float y2 = (float)YD[2 * MAXTAPS + p + R1], y3 = (float)YD[3 * MAXTAPS + p + R1];
        g[2] = fmaf(h2, Y0[p + R1], g[2]);
        g[3] = fmaf(h1, Y1[p + R1], g[3]);
        g[4] = fmaf(h0, y2, g[4]);
        g[5] = fmaf(h2, Y1[p + R1], g[5]);
        g[6] = fmaf(h1, y2, g[6]);
        g[7] = fmaf(h3, Y0[p + R1], g[7]);
        g[8] = fmaf(h0, y3, g[8]);
    }
    const float fxx = g[2], fxy = g[3], fyy = g[4], fxxy = g[5], fxyy = g[6], fxxx = g[7], fyyy = g[8];
    float tx = fx * (2 * fxx * fxx + 2 * fxy * fxy) + fy * (2 * fxx * fxy + 2 * fyy * fxy)
             + 2 * fx * fy * fxxy + fy * fy * fxyy + fx * fx * fxxx;
    float ty = fx * (2 * fxx * fxy + 2 * fyy * fxy) + fy * (2 * fyy * fyy + 2 * fxy * fxy)
             + 2 * fx * fy * fxyy + fx * fx * fxxy + fy * fy * fyyy;
    const float m = sqrtf(tx * tx + ty * ty);
    tx /= m;
    ty /= m;
    Ix[o] = fx;
    Iy[o] = fy;
    mag[o] = sqrtf(fx * fx + fy * fy);
    orient[o] = atan2f(tx, -ty);
}

__global__ __launch_bounds__(256) void toed_nms_maps_kernel(int IH, int IW, const float *Ix, const float *Iy,
                                                            const float *mag, float *px, float *py, float thresh,
                                                            int border) {
    const long long o = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (o >= (long long)IH * IW) return;
    const int i = (int)(o / IW), j = (int)(o % IW);
    float ox = 0.f, oy = 0.f;
    const float f = mag[o];
    if (i >= border && i < IH - border && j >= border && j < IW - border && f > thresh) {
        const float gx = Ix[o], gy = Iy[o];
        if (!(fabsf(gx) < 10e-6f && fabsf(gy) < 10e-6f)) {
            auto Mf = [&](int dr, int dc) { return mag[o + (long long)dr * IW + dc]; };
            const float nx = gx / f, ny = gy / f;
            float sl, fp, fm;
            if (octant<float>(gx, gy, nx, ny, Mf, sl, fp, fm) &&
                ((f > fm && f > fp) || (f > fm && f >= fp) || (f >= fm && f > fp))) {
                auto Md = [&](int dr, int dc) { return (double)mag[o + (long long)dr * IW + dc]; };
                const double fd = f, nxd = (double)gx / fd, nyd = (double)gy / fd;
                double sld, fpd, fmd;
                octant<double>((double)gx, (double)gy, nxd, nyd, Md, sld, fpd, fmd);
                const double s = sqrt(1.0 + sld * sld);
                const double A = (fmd + fpd - 2.0 * fd) / (2.0 * s * s);
                const double Bc = (fpd - fmd) / (2.0 * s);
                const double ss = -Bc / (2.0 * A);
                if (fabs(ss) <= 1.4142135623730951) {
                    ox = (float)(j + ss * nxd);
                    oy = (float)(i + ss * nyd);
                }
            }
        }
    }
    px[o] = ox;
    py[o] = oy;
}


// FP64 form of the unfused launch layer (the reference's T = double overloads
// of gpu_convolve / gpu_nms, gpu_kernels.hpp:35-45,56-63): the same separable
// sums on the FP64 taps, one thread per interp pixel, and the NMS decided
// wholly in FP64 (cpu_toed.cpp:400-494).
__global__ __launch_bounds__(256) void toed_planes_f64_kernel(int H, int W, const double *img, double *Ix,
                                                              double *Iy, double *mag, double *orient,
                                                              const TapSet *tp) {
    const long long o = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long IW = 2LL * W;
    if (o >= 4LL * H * W) return;
    const int R = (int)(o / IW), C = (int)(o % IW);
    const int i = R >> 1, j = C >> 1, a = R & 1, b = C & 1;
    const int R1 = tp->R1, rad = (a | b) ? tp->R1 : tp->R0;
    const double(*X)[MAXTAPS] = b ? tp->ds : tp->du;
    const double(*Y)[MAXTAPS] = a ? tp->ds : tp->du;
    double g[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int p = -rad; p <= rad; ++p) {
        const int ii = i - p;
        if (ii < 0 || ii >= H) continue;
        double h0 = 0, h1 = 0, h2 = 0, h3 = 0;
        for (int q = -rad; q <= rad; ++q) {
            const int jj = j - q;
            if (jj < 0 || jj >= W) continue;
            const double v = img[(long long)ii * W + jj];
            const int t = q + R1;
            h0 = fma(v, X[0][t], h0);
            h1 = fma(v, X[1][t], h1);
            h2 = fma(v, X[2][t], h2);
            h3 = fma(v, X[3][t], h3);
        }
        const int t = p + R1;
        g[0] = fma(h1, Y[0][t], g[0]);   // fx
        g[1] = fma(h0, Y[1][t], g[1]);   // fy
        g[2] = fma(h2, Y[0][t], g[2]);   // fxx
        g[3] = fma(h1, Y[1][t], g[3]);   // fxy
        g[4] = fma(h0, Y[2][t], g[4]);   // fyy
        g[5] = fma(h2, Y[1][t], g[5]);   // fxxy
        g[6] = fma(h1, Y[2][t], g[6]);   // fxyy
        g[7] = fma(h3, Y[0][t], g[7]);   // fxxx
        g[8] = fma(h0, Y[3][t], g[8]);   // fyyy
    }
    const double fx = g[0], fy = g[1], fxx = g[2], fxy = g[3], fyy = g[4];
    const double fxxy = g[5], fxyy = g[6], fxxx = g[7], fyyy = g[8];
    double tx = fx * (2 * fxx * fxx + 2 * fxy * fxy) + fy * (2 * fxx * fxy + 2 * fyy * fxy)
              + 2 * fx * fy * fxxy + fy * fy * fxyy + fx * fx * fxxx;
    double ty = fx * (2 * fxx * fxy + 2 * fyy * fxy) + fy * (2 * fyy * fyy + 2 * fxy * fxy)
              + 2 * fx * fy * fxyy + fx * fx * fxxy + fy * fy * fyyy;
    const double m = sqrt(tx * tx + ty * ty);
    tx /= m;
    ty /= m;
    Ix[o] = fx;
    Iy[o] = fy;
    mag[o] = sqrt(fx * fx + fy * fy);
    orient[o] = atan2(tx, -ty);
}

__global__ __launch_bounds__(256) void toed_nms_maps_f64_kernel(int IH, int IW, const double *Ix, const double *Iy,
                                                                const double *mag, double *px, double *py,
                                                                double thresh, int border) {
    const long long o = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (o >= (long long)IH * IW) return;
    const int i = (int)(o / IW), j = (int)(o % IW);
    double ox = 0.0, oy = 0.0;
    const double f = mag[o];
    if (i >= border && i < IH - border && j >= border && j < IW - border && f > thresh) {
        const double gx = Ix[o], gy = Iy[o];
        if (!(fabs(gx) < 10e-6 && fabs(gy) < 10e-6)) {
            auto M = [&](int dr, int dc) { return mag[o + (long long)dr * IW + dc]; };
            const double nx = gx / f, ny = gy / f;
            double sl, fp, fm;
            if (octant<double>(gx, gy, nx, ny, M, sl, fp, fm) &&
                ((f > fm && f > fp) || (f > fm && f >= fp) || (f >= fm && f > fp))) {
                const double s = sqrt(1.0 + sl * sl);
                const double A = (fm + fp - 2.0 * f) / (2.0 * s * s);
                const double Bc = (fp - fm) / (2.0 * s);
                const double ss = -Bc / (2.0 * A);
                if (fabs(ss) <= 1.4142135623730951) {
                    ox = j + ss * nx;
                    oy = i + ss * ny;
                }
            }
        }
    }
    px[o] = ox;
    py[o] = oy;
}

}  // namespace toed

// ===========================================================================
// Host side
// ===========================================================================
using namespace toed;

// tiles of the images with one window radius (contiguous in the tile order)
struct TileGroup {
    int R1 = 0, tile_begin = 0, tile_end = 0, grid_persistent = 0;
    int occ = 1;                // tile-kernel workgroups that fit on one CU
};

struct toed_ctx {
    int device = 0;
    hipStream_t own_stream = nullptr;
    hipStream_t stream = nullptr;
    int tile_rows = 16;
    float stage_frac = 1.0f;
    int orient_mode = 0;       // 1: orientation at the subpixel location
    // tap sets by sigma
    std::vector<float> tap_sigmas;
    std::vector<TapSet *> tap_dev;
    // single-image scratch (grow-only)
    int *d_counts = nullptr, *d_offsets = nullptr, *d_chunks = nullptr, *d_tilecnt = nullptr;
    float4 *d_srec = nullptr;
    uint32_t *d_skey = nullptr;
    size_t n_counts = 0, n_offsets = 0, n_chunks = 0, n_tiles = 0, n_srec = 0, n_skey = 0;
    int *d_scalar = nullptr;   // [0] count, [1] flags for *_sync
    hipEvent_t ev[3] = {nullptr, nullptr, nullptr};
    float ms_tiles = 0.f, ms_final = 0.f;
};

struct toed_batch {
    toed_ctx *ctx = nullptr;
    int n = 0;
    std::vector<ImageJob> jobs;
    ImageJob *d_jobs = nullptr;
    int *d_counts = nullptr, *d_offsets = nullptr, *d_chunks = nullptr, *d_tilecnt = nullptr, *d_ticket = nullptr;
    float4 *d_srec = nullptr;
    uint32_t *d_skey = nullptr;
    long long ntiles = 0, ncounts = 0, nchunks = 0, nstaging = 0;
    int in_type = 0, tile_rows = 16, persistent = 0;
    int dev_cu = 1, wg_per_cu = 0;     // persistent grid: workgroups per CU (0: as many as fit)
    bool qad = false;                  // orientation at the subpixel location (all jobs)
    std::vector<int> order;
    std::vector<TileGroup> groups;
    GradOut *d_gouts = nullptr;
    GradOut *h_gouts = nullptr;        // pinned staging of the last upload of d_gouts
    std::vector<GradOut> last_gouts;   // what d_gouts holds (an unchanged table is not uploaded again)
    hipEvent_t gouts_ev = nullptr;     // recorded after the last upload (guards h_gouts)
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    int *g_counts = nullptr, *g_flags = nullptr;
    bool dirty = false;                // the job table changed on the host (images re-pointed, compact output)
    bool uploaded = false;             // the device job table holds the host's
    long long *d_cbase = nullptr;      // compact output: each image's first row (caller's order)
    long long ccap = 0;
};

static std::mutex g_mu;

// Scratch of the allocation-free entry point (toed_detect_noalloc): one
// statically sized array per GPU, loaded with the code object and held until
// the process ends.  The host never passes `scratch` itself to a kernel (in
// host code the name is a stand-in, not a device address): each device's
// address comes from hipGetSymbolAddress once, at context creation, and is
// kept in g_noalloc_ptr.  Its size holds a 2160 x 3840 frame at 16-row tiles
// (taps, counts and their scan, per-tile counts, 64 staging records per tile
// and phase: 87.4 MB).
constexpr size_t NOALLOC_BYTES = TOED_NOALLOC_SCRATCH_BYTES;
constexpr int NOALLOC_STAGE_CAP = 64;
constexpr int NOALLOC_MAX_DEVICES = 64;
static __device__ __align__(256) unsigned char scratch[NOALLOC_BYTES];
static unsigned char *g_noalloc_ptr[NOALLOC_MAX_DEVICES];

#define HIPCHK(x)                                              \
    do {                                                       \
        hipError_t e_ = (x);                                   \
        if (e_ != hipSuccess) {                                \
            std::fprintf(stderr, "toed: %s failed: %s (%s:%d)\n", #x, hipGetErrorString(e_), __FILE__, __LINE__); \
            return TOED_ERR_HIP;                               \
        }                                                      \
    } while (0)

static int radius_for(float sigma, int *R0, int *R1) {
    if (!(sigma > 0.f) || !std::isfinite(sigma)) return TOED_ERR_SIGMA;
    const int r0 = (int)std::ceil(4.0 * (double)sigma - 1e-9);
    const int r1 = r0 + 1;
    if (r1 != 5 && r1 != 7 && r1 != 9) return TOED_ERR_SIGMA;
    *R0 = r0;
    *R1 = r1;
    return TOED_OK;
}

static void fill_taps(double sigma, int R0, int R1, TapSet *t) {
    std::memset(t, 0, sizeof(*t));
    t->R0 = R0;
    t->R1 = R1;
    const double s2 = sigma * sigma, norm = std::sqrt(2.0 * M_PI) * sigma;
    for (int sh = 0; sh < 2; ++sh) {
        for (int k = 0; k <= 2 * R1; ++k) {
            const double x = (double)(k - R1) + (sh ? 0.5 : 0.0);
            const double g = std::exp(-x * x / (2.0 * s2)) / norm;
            const double v[4] = {g, -x * g / s2, (x * x - s2) * g / (s2 * s2), x * (3.0 * s2 - x * x) * g / (s2 * s2 * s2)};
            // d/dsigma: f = P(x, s) g with dg/ds = g (x^2/s^3 - 1/s)
            const double s1 = sigma, dg = g * (x * x / (s2 * s1) - 1.0 / s1);
            const double P[4] = {1.0, -x / s2, (x * x - s2) / (s2 * s2), x * (3.0 * s2 - x * x) / (s2 * s2 * s2)};
            const double dP[4] = {0.0, 2.0 * x / (s2 * s1), -4.0 * x * x / (s2 * s2 * s1) + 2.0 / (s2 * s1),
                                  -12.0 * x / (s2 * s2 * s1) + 6.0 * x * x * x / (s2 * s2 * s2 * s1)};
            for (int f = 0; f < 4; ++f) {
                const double dv = dP[f] * g + P[f] * dg;
                if (sh) t->dss[f][k] = dv; else t->dsu[f][k] = dv;
                if (sh) t->ds[f][k] = v[f]; else t->du[f][k] = v[f];
                if (f < 2) { if (sh) t->fs[f][k] = (float)v[f]; else t->fu[f][k] = (float)v[f]; }
            }
        }
    }
}

static int get_taps(toed_ctx *c, float sigma, const TapSet **out) {
    for (size_t k = 0; k < c->tap_sigmas.size(); ++k)
        if (c->tap_sigmas[k] == sigma) { *out = c->tap_dev[k]; return TOED_OK; }
    int R0, R1;
    int st = radius_for(sigma, &R0, &R1);
    if (st) return st;
    TapSet h;
    fill_taps(sigma, R0, R1, &h);
    TapSet *d = nullptr;
    if (hipMalloc(&d, sizeof(TapSet)) != hipSuccess) return TOED_ERR_ALLOC;
    HIPCHK(hipMemcpy(d, &h, sizeof(TapSet), hipMemcpyHostToDevice));
    c->tap_sigmas.push_back(sigma);
    c->tap_dev.push_back(d);
    *out = d;
    return TOED_OK;
}

static int check_image(const toed_image_t *im) {
    if (!im || !im->d_img || im->H <= 0 || im->W <= 0 || im->pitch < im->W) return TOED_ERR_ARG;
    if (im->in_type != TOED_U8 && im->in_type != TOED_F32) return TOED_ERR_ARG;
    if (im->border < 1 || im->capacity < 0 || (im->capacity > 0 && !im->d_edges)) return TOED_ERR_ARG;
    if ((long long)im->H * im->W > (1LL << 28)) return TOED_ERR_ARG;   // 2H*2W must fit int32 indices
    return TOED_OK;
}

// fills geometry fields of a job; returns number of tiles/count entries/chunks
static void job_geometry(ImageJob &J, int TH, int stage_cap) {
    J.ntr = ceil_div(J.H, TH);
    J.ntc = ceil_div(J.W, TW);
    const long long ncount = (long long)(2 * J.H) * J.ntc;
    J.nchunks = (int)((ncount + CHUNK - 1) / CHUNK);
    J.stage_cap = stage_cap;
}

static int stage_cap_for(float frac, int TH) {
    int cap = (int)std::ceil((double)frac * TH * TW);
    if (cap < 1) cap = 1;
    if (cap > TH * TW) cap = TH * TW;
    return cap;
}

template <typename F> static void for_tile_rows(int TH, F f) {
    if (TH == 8) f(std::integral_constant<int, 8>());
    else f(std::integral_constant<int, 16>());
}

template <typename F> static void for_radius(int R1, F f) {
    if (R1 == 5) f(std::integral_constant<int, 5>());
    else if (R1 == 7) f(std::integral_constant<int, 7>());
    else f(std::integral_constant<int, 9>());
}

static int group_of(int R1) { return R1 == 5 ? 0 : (R1 == 7 ? 1 : 2); }


template <typename IN_T, int T, int R>
static const void *tiles_kernel_ptr() { return (const void *)toed_tiles_kernel<IN_T, T, R>; }

// enqueue the kernels for jobs (device table) or job0 (single image)
static int enqueue(hipStream_t s, int in_type, int TH, const ImageJob *d_jobs, int njobs, const ImageJob &job0,
                   const TileGroup *groups, int ngroups, long long ntiles, long long nchunks, int *counts,
                   int *offsets, int *chunks, StageBuf staging, int *tilecnt, int *tickets, hipEvent_t *ev,
                   bool qad, const int *ccounts = nullptr, int *cflags = nullptr, long long ccap = 0,
                   long long *cbase = nullptr) {
    if (ntiles <= 0) return TOED_OK;
    int *tlist = tilecnt + 4 * ntiles;      // count, then the tiles with edgels
    hipLaunchKernelGGL(toed_reset_kernel, dim3(ceil_div(njobs, 256)), dim3(256), 0, s, d_jobs, njobs, job0, tlist);
    if (ev) HIPCHK(hipEventRecord(ev[0], s));
    for (int gi = 0; gi < ngroups; ++gi) {
        const TileGroup &g = groups[gi];
        if (g.tile_end <= g.tile_begin) continue;
        int *ticket = tickets ? tickets + group_of(g.R1) : nullptr;
        const dim3 grid(ticket ? (unsigned)g.grid_persistent : (unsigned)(g.tile_end - g.tile_begin));
        for_tile_rows(TH, [&](auto THc) {
            constexpr int T = decltype(THc)::value;
            for_radius(g.R1, [&](auto Rc) {
                constexpr int R = decltype(Rc)::value;
                if (in_type == TOED_U8)
                    hipLaunchKernelGGL((toed_tiles_kernel<uint8_t, T, R>), grid, dim3(NT), 0, s, d_jobs, njobs, job0,
                                       counts, staging, tilecnt, tlist, g.tile_begin, g.tile_end, ticket);
                else
                    hipLaunchKernelGGL((toed_tiles_kernel<float, T, R>), grid, dim3(NT), 0, s, d_jobs, njobs, job0,
                                       counts, staging, tilecnt, tlist, g.tile_begin, g.tile_end, ticket);
            });
        });
    }
    if (qad)
        hipLaunchKernelGGL(toed_qad_kernel, dim3((unsigned)ntiles), dim3(NT), 0, s, d_jobs, njobs, job0, staging,
                           tilecnt);
    if (ev) HIPCHK(hipEventRecord(ev[1], s));
    hipLaunchKernelGGL(toed_scan_chunks, dim3((unsigned)nchunks), dim3(256), 0, s, d_jobs, njobs, job0, counts,
                       offsets, chunks);
    hipLaunchKernelGGL(toed_scan_images, dim3(njobs), dim3(1024), 0, s, d_jobs, njobs, job0, chunks);
    if (cbase)
        hipLaunchKernelGGL(toed_compact_base_kernel, dim3(1), dim3(1024), 0, s, ccounts, cflags, njobs, ccap, cbase);
    // finalize: a grid that walks the tiles with edgels
    hipLaunchKernelGGL(toed_finalize_kernel, dim3((unsigned)std::min<long long>(ntiles, FIN_GRID)), dim3(FIN_NT), 0, s,
                       d_jobs, njobs, job0, TH, offsets, chunks, staging, tilecnt, tlist, tickets);
    if (ev) HIPCHK(hipEventRecord(ev[2], s));
    HIPCHK(hipGetLastError());
    return TOED_OK;
}

template <typename T> static int grow(T **p, size_t *have, size_t need) {
    if (*have >= need && *p) return TOED_OK;
    if (*p) (void)hipFree(*p);
    *p = nullptr;
    *have = 0;
    if (hipMalloc((void **)p, std::max<size_t>(need, 1) * sizeof(T)) != hipSuccess) { *p = nullptr; return TOED_ERR_ALLOC; }
    *have = need;
    return TOED_OK;
}

extern "C" {

int toed_version(void) { return 100; }

int toed_sigma_radius(float sigma, int *R0, int *R1) {
    if (!R0 || !R1) return TOED_ERR_ARG;
    return radius_for(sigma, R0, R1);
}

int toed_create(int device, toed_ctx **out) {
    if (!out) return TOED_ERR_ARG;
    *out = nullptr;
    int ndev = 0;
    HIPCHK(hipGetDeviceCount(&ndev));
    if (device < 0 || device >= ndev) return TOED_ERR_ARG;
    HIPCHK(hipSetDevice(device));
    if (device < NOALLOC_MAX_DEVICES) {
        // the allocation-free entry point's array on this device: its address,
        // once per device, with the device current
        std::lock_guard<std::mutex> lk(g_mu);
        if (!g_noalloc_ptr[device]) {
            void *p = nullptr;
            HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
            g_noalloc_ptr[device] = (unsigned char *)p;
        }
    }
    toed_ctx *c = new toed_ctx();
    c->device = device;
    // blocking stream: orders with the legacy default stream, so work launched
    // here (e.g. graph replays) is ordered with a caller on the null stream
    if (hipStreamCreateWithFlags(&c->own_stream, hipStreamDefault) != hipSuccess) { delete c; return TOED_ERR_HIP; }
    c->stream = c->own_stream;
    for (auto &e : c->ev) HIPCHK(hipEventCreate(&e));
    if (hipMalloc(&c->d_scalar, 2 * sizeof(int)) != hipSuccess) { delete c; return TOED_ERR_ALLOC; }
    *out = c;
    return TOED_OK;
}

int toed_destroy(toed_ctx *c) {
    if (!c) return TOED_ERR_ARG;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    for (auto *t : c->tap_dev) (void)hipFree(t);
    (void)hipFree(c->d_counts); (void)hipFree(c->d_offsets); (void)hipFree(c->d_chunks);
    (void)hipFree(c->d_tilecnt); (void)hipFree(c->d_srec); (void)hipFree(c->d_skey); (void)hipFree(c->d_scalar);
    for (auto &e : c->ev) if (e) (void)hipEventDestroy(e);
    if (c->own_stream) (void)hipStreamDestroy(c->own_stream);
    delete c;
    return TOED_OK;
}

int toed_set_stream(toed_ctx *c, void *s) {
    if (!c) return TOED_ERR_ARG;
    c->stream = (hipStream_t)s;
    return TOED_OK;
}

int toed_use_own_stream(toed_ctx *c) {
    if (!c) return TOED_ERR_ARG;
    c->stream = c->own_stream;
    return TOED_OK;
}

void *toed_get_stream(toed_ctx *c) { return c ? (void *)c->stream : nullptr; }

int toed_get_device(toed_ctx *c) { return c ? c->device : -1; }

int toed_set_tile_rows(toed_ctx *c, int tr) {
    if (!c) return TOED_ERR_ARG;
    if (tr == 0) tr = 16;
    if (tr != 8 && tr != 16) return TOED_ERR_ARG;
    c->tile_rows = tr;
    return TOED_OK;
}

int toed_set_orientation_mode(toed_ctx *c, int mode) {
    if (!c || (mode != 0 && mode != 1)) return TOED_ERR_ARG;
    c->orient_mode = mode;
    return TOED_OK;
}

int toed_set_staging_fraction(toed_ctx *c, float f) {
    if (!c || !(f > 0.f) || f > 1.f) return TOED_ERR_ARG;
    c->stage_frac = f;
    return TOED_OK;
}

int toed_last_times(toed_ctx *c, float *a, float *b) {
    if (!c) return TOED_ERR_ARG;
    if (a) *a = c->ms_tiles;
    if (b) *b = c->ms_final;
    return TOED_OK;
}

static int detect_impl(toed_ctx *c, const toed_image_t *im, int *d_count, int *d_flags, bool timed) {
    if (!c || !d_count) return TOED_ERR_ARG;
    int st = check_image(im);
    if (st) return st;
    HIPCHK(hipSetDevice(c->device));
    const TapSet *taps = nullptr;
    if ((st = get_taps(c, im->sigma, &taps))) return st;
    const int TH = c->tile_rows;
    ImageJob J;
    std::memset(&J, 0, sizeof J);
    J.img = im->d_img; J.out = im->d_edges; J.count_out = d_count; J.flags_out = d_flags; J.taps = taps;
    J.index_out = im->d_index;
    J.capacity = im->capacity; J.in_type = im->in_type; J.H = im->H; J.W = im->W; J.pitch = im->pitch;
    radius_for(im->sigma, &J.R0, &J.R1);
    J.border = im->border; J.thresh = im->thresh;
    J.qad = c->orient_mode;
    job_geometry(J, TH, stage_cap_for(c->stage_frac, TH));
    const long long ntiles = (long long)J.ntr * J.ntc;
    const long long ncounts = (long long)J.nchunks * CHUNK;
    if ((st = grow(&c->d_counts, &c->n_counts, (size_t)ncounts))) return st;
    if ((st = grow(&c->d_offsets, &c->n_offsets, (size_t)ncounts))) return st;
    if ((st = grow(&c->d_chunks, &c->n_chunks, (size_t)J.nchunks))) return st;
    if ((st = grow(&c->d_tilecnt, &c->n_tiles, (size_t)ntiles * 5 + 1))) return st;    // + the tile list
    if ((st = grow(&c->d_srec, &c->n_srec, (size_t)ntiles * 4 * J.stage_cap))) return st;
    if ((st = grow(&c->d_skey, &c->n_skey, (size_t)ntiles * 4 * J.stage_cap))) return st;
    J.counts_base = 0; J.chunk_base = 0; J.stage_base = 0; J.tile_begin = 0;
    TileGroup g;
    g.R1 = J.R1; g.tile_begin = 0; g.tile_end = (int)ntiles;
    return enqueue(c->stream, J.in_type, TH, nullptr, 1, J, &g, 1, ntiles, J.nchunks,
                   c->d_counts, c->d_offsets, c->d_chunks, StageBuf{c->d_srec, c->d_skey}, c->d_tilecnt, nullptr,
                   timed ? c->ev : nullptr, J.qad != 0);
}

int toed_detect(toed_ctx *c, const toed_image_t *im, int *d_count, int *d_flags) {
    return detect_impl(c, im, d_count, d_flags, false);
}

int toed_detect_sync(toed_ctx *c, const toed_image_t *im, long long *n_out) {
    if (!c || !n_out) return TOED_ERR_ARG;
    int st = detect_impl(c, im, c->d_scalar, c->d_scalar + 1, true);
    if (st) return st;
    int h[2] = {0, 0};
    HIPCHK(hipMemcpyAsync(h, c->d_scalar, 2 * sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    HIPCHK(hipEventElapsedTime(&c->ms_tiles, c->ev[0], c->ev[1]));
    HIPCHK(hipEventElapsedTime(&c->ms_final, c->ev[1], c->ev[2]));
    *n_out = h[0];
    if (h[1] & TOED_FLAG_STAGING) return TOED_ERR_STAGING;
    if (h[1] & TOED_FLAG_CAPACITY) return TOED_ERR_CAPACITY;
    return TOED_OK;
}

// toed_detect with every scratch buffer carved from the device's static array
// (no hipMalloc, no copy): taps by a kernel with by-value arguments, then the
// same kernels as toed_detect.  Nothing in the array has to start at zero: the
// tile kernel writes every count and per-tile count the scans read, staging
// records are written before they are read, and the flags are reset by a
// kernel.  Overlapping calls on one GPU share the array and race.
int toed_detect_noalloc(toed_ctx *c, const toed_image_t *im, int *d_count, int *d_flags) {
    if (!c || !d_count) return TOED_ERR_ARG;
    int st = check_image(im);
    if (st) return st;
    if (c->device < 0 || c->device >= NOALLOC_MAX_DEVICES || !g_noalloc_ptr[c->device]) return TOED_ERR_ARG;
    HIPCHK(hipSetDevice(c->device));
    int R0, R1;
    if ((st = radius_for(im->sigma, &R0, &R1))) return st;
    const int TH = c->tile_rows;
    ImageJob J;
    std::memset(&J, 0, sizeof J);
    J.img = im->d_img; J.out = im->d_edges; J.count_out = d_count; J.flags_out = d_flags;
    J.index_out = im->d_index;
    J.capacity = im->capacity; J.in_type = im->in_type; J.H = im->H; J.W = im->W; J.pitch = im->pitch;
    J.R0 = R0; J.R1 = R1;
    J.border = im->border; J.thresh = im->thresh;
    J.qad = c->orient_mode;
    job_geometry(J, TH, std::min(NOALLOC_STAGE_CAP, stage_cap_for(c->stage_frac, TH)));
    const long long ntiles = (long long)J.ntr * J.ntc;
    const long long ncounts = (long long)J.nchunks * CHUNK;
    // carve the array: 256-byte aligned pieces
    auto up = [](size_t b) { return (b + 255) & ~size_t(255); };
    const size_t o_taps = 0;
    const size_t o_counts = o_taps + up(sizeof(TapSet));
    const size_t o_offsets = o_counts + up(sizeof(int) * ncounts);
    const size_t o_chunks = o_offsets + up(sizeof(int) * ncounts);
    const size_t o_tilecnt = o_chunks + up(sizeof(int) * J.nchunks);
    const size_t o_srec = o_tilecnt + up(sizeof(int) * (ntiles * 5 + 1));   // per-tile counts, the tile list
    const size_t o_skey = o_srec + up(sizeof(float4) * ntiles * 4 * J.stage_cap);
    const size_t end = o_skey + up(sizeof(uint32_t) * ntiles * 4 * J.stage_cap);
    if (end > NOALLOC_BYTES) {
        std::fprintf(stderr, "toed_detect_noalloc: a %d x %d image needs %zu bytes of scratch, more than the %zu "
                             "the library holds\n", J.H, J.W, end, NOALLOC_BYTES);
        return TOED_ERR_ARG;
    }
    unsigned char *base = g_noalloc_ptr[c->device];
    TapSet h;
    fill_taps(im->sigma, R0, R1, &h);
    TapArgs a;
    a.R0 = R0; a.R1 = R1;
    std::memcpy(a.fu, h.fu, sizeof a.fu); std::memcpy(a.fs, h.fs, sizeof a.fs);
    std::memcpy(a.du, h.du, sizeof a.du); std::memcpy(a.ds, h.ds, sizeof a.ds);
    TapSet *taps = (TapSet *)(base + o_taps);
    hipLaunchKernelGGL(toed_put_taps_kernel, dim3(1), dim3(256), 0, c->stream, taps, a);
    J.taps = taps;
    J.counts_base = 0; J.chunk_base = 0; J.stage_base = 0; J.tile_begin = 0;
    TileGroup g;
    g.R1 = J.R1; g.tile_begin = 0; g.tile_end = (int)ntiles;
    return enqueue(c->stream, J.in_type, TH, nullptr, 1, J, &g, 1, ntiles, J.nchunks,
                   (int *)(base + o_counts), (int *)(base + o_offsets), (int *)(base + o_chunks),
                   StageBuf{(float4 *)(base + o_srec), (uint32_t *)(base + o_skey)}, (int *)(base + o_tilecnt),
                   nullptr, nullptr, J.qad != 0);
}

// persistent grid of each radius group: wg_per_cu workgroups on every CU
// (at most what fits), never more than the group's tiles
static void set_grids(toed_batch *b) {
    for (TileGroup &g : b->groups) {
        const int per_cu = b->wg_per_cu > 0 ? std::min(b->wg_per_cu, g.occ) : g.occ;
        g.grid_persistent = std::max(1, std::min(b->dev_cu * per_cu, g.tile_end - g.tile_begin));
    }
}

int toed_batch_plan(toed_ctx *c, int n, const toed_image_t *imgs, toed_batch **out) {
    if (!c || n <= 0 || !imgs || !out) return TOED_ERR_ARG;
    *out = nullptr;
    HIPCHK(hipSetDevice(c->device));
    const int in_type = imgs[0].in_type;
    const int TH = c->tile_rows;
    const int scap = stage_cap_for(c->stage_frac, TH);
    toed_batch *b = new toed_batch();
    b->ctx = c; b->n = n; b->in_type = in_type; b->tile_rows = TH; b->qad = c->orient_mode != 0;
    b->jobs.resize(n);
    // jobs ordered by window radius so each radius' tiles are contiguous and
    // run in one specialised kernel; order[k] is the caller's image index
    b->order.resize(n);
    for (int m = 0; m < n; ++m) b->order[m] = m;
    std::vector<int> r1s(n, 0);
    for (int m = 0; m < n; ++m) {
        int r0, r1;
        if (radius_for(imgs[m].sigma, &r0, &r1)) { delete b; return TOED_ERR_SIGMA; }
        r1s[m] = r1;
    }
    std::stable_sort(b->order.begin(), b->order.end(), [&](int x, int y) { return r1s[x] < r1s[y]; });
    long long tiles = 0, chunks = 0;
    for (int k = 0; k < n; ++k) {
        const int m = b->order[k];
        const toed_image_t *im = &imgs[m];
        int st = check_image(im);
        if (!st && im->in_type != in_type) st = TOED_ERR_ARG;
        const TapSet *taps = nullptr;
        if (!st) st = get_taps(c, im->sigma, &taps);
        if (st) { delete b; return st; }
        ImageJob &J = b->jobs[k];
        std::memset(&J, 0, sizeof J);
        J.img = im->d_img; J.out = im->d_edges; J.taps = taps; J.capacity = im->capacity;
        J.index_out = im->d_index;
        J.in_type = in_type; J.H = im->H; J.W = im->W; J.pitch = im->pitch;
        radius_for(im->sigma, &J.R0, &J.R1);
        J.border = im->border; J.thresh = im->thresh;
        J.qad = c->orient_mode;
        job_geometry(J, TH, scap);
        J.tile_begin = (int)tiles;
        J.chunk_base = chunks;
        J.counts_base = chunks * CHUNK;
        J.stage_base = tiles * 4 * scap;
        tiles += (long long)J.ntr * J.ntc;
        chunks += J.nchunks;
        if (b->groups.empty() || b->groups.back().R1 != J.R1) {
            TileGroup g;
            g.R1 = J.R1;
            g.tile_begin = J.tile_begin;
            b->groups.push_back(g);
        }
        b->groups.back().tile_end = (int)tiles;
    }
    if (tiles >= (1LL << 31)) { delete b; return TOED_ERR_ARG; }
    b->ntiles = tiles; b->nchunks = chunks; b->ncounts = chunks * CHUNK; b->nstaging = tiles * 4 * scap;
    bool ok = hipMalloc(&b->d_jobs, sizeof(ImageJob) * n) == hipSuccess &&
              hipMalloc(&b->d_counts, sizeof(int) * b->ncounts) == hipSuccess &&
              hipMalloc(&b->d_offsets, sizeof(int) * b->ncounts) == hipSuccess &&
              hipMalloc(&b->d_chunks, sizeof(int) * std::max<long long>(chunks, 1)) == hipSuccess &&
              hipMalloc(&b->d_tilecnt, sizeof(int) * (tiles * 5 + 1)) == hipSuccess &&   // + the tile list
              hipMalloc(&b->d_ticket, sizeof(int) * NGROUPS) == hipSuccess &&
              hipMalloc(&b->d_srec, sizeof(float4) * std::max<long long>(b->nstaging, 1)) == hipSuccess &&
              hipMalloc(&b->d_skey, sizeof(uint32_t) * std::max<long long>(b->nstaging, 1)) == hipSuccess;
    if (!ok) { toed_batch_destroy(b); return TOED_ERR_ALLOC; }
    if (hipMemset(b->d_ticket, 0, sizeof(int) * NGROUPS) != hipSuccess) { toed_batch_destroy(b); return TOED_ERR_HIP; }
    int dev_cu = 0;
    (void)hipDeviceGetAttribute(&dev_cu, hipDeviceAttributeMultiprocessorCount, c->device);
    for (TileGroup &g : b->groups) {
        int occ = 0;
        for_tile_rows(TH, [&](auto THc) {
            constexpr int T = decltype(THc)::value;
            for_radius(g.R1, [&](auto Rc) {
                constexpr int R = decltype(Rc)::value;
                if (in_type == TOED_U8)
                    (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, toed_tiles_kernel<uint8_t, T, R>, NT, 0);
                else
                    (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, toed_tiles_kernel<float, T, R>, NT, 0);
            });
        });
        g.occ = std::max(occ, 1);
    }
    b->dev_cu = std::max(dev_cu, 1);
    set_grids(b);
    *out = b;
    return TOED_OK;
}

int toed_batch_set_persistent(toed_batch *b, int p) {
    if (!b) return TOED_ERR_ARG;
    b->persistent = p ? 1 : 0;
    if (b->gexec) { (void)hipGraphExecDestroy(b->gexec); b->gexec = nullptr; }
    if (b->graph) { (void)hipGraphDestroy(b->graph); b->graph = nullptr; }
    return TOED_OK;
}

int toed_batch_set_grid(toed_batch *b, int wg_per_cu) {
    if (!b || wg_per_cu < 0) return TOED_ERR_ARG;
    b->wg_per_cu = wg_per_cu;
    set_grids(b);
    if (b->gexec) { (void)hipGraphExecDestroy(b->gexec); b->gexec = nullptr; }
    if (b->graph) { (void)hipGraphDestroy(b->graph); b->graph = nullptr; }
    return TOED_OK;
}

int toed_batch_set_images(toed_batch *b, const void *const *d_imgs) {
    if (!b || !d_imgs) return TOED_ERR_ARG;
    for (int k = 0; k < b->n; ++k)
        if (!d_imgs[b->order[k]]) return TOED_ERR_ARG;
    for (int k = 0; k < b->n; ++k) b->jobs[k].img = d_imgs[b->order[k]];
    b->dirty = true;
    return TOED_OK;
}

int toed_batch_set_compact_output(toed_batch *b, float *d_edges, int *d_index, long long capacity) {
    if (!b || !d_edges || capacity <= 0) return TOED_ERR_ARG;
    HIPCHK(hipSetDevice(b->ctx->device));
    if (!b->d_cbase && hipMalloc(&b->d_cbase, sizeof(long long) * b->n) != hipSuccess) {
        b->d_cbase = nullptr;
        return TOED_ERR_ALLOC;
    }
    for (int k = 0; k < b->n; ++k) {
        ImageJob &J = b->jobs[k];
        J.out = d_edges;
        J.index_out = d_index;
        J.capacity = capacity;
        J.cbase = b->d_cbase + b->order[k];
    }
    b->ccap = capacity;
    b->dirty = true;
    return TOED_OK;
}

static int batch_upload(toed_batch *b, int *d_counts, int *d_flags) {
    for (int k = 0; k < b->n; ++k) {
        const int m = b->order[k];
        b->jobs[k].count_out = d_counts ? d_counts + m : nullptr;
        b->jobs[k].flags_out = d_flags ? d_flags + m : nullptr;
    }
    HIPCHK(hipMemcpy(b->d_jobs, b->jobs.data(), sizeof(ImageJob) * b->n, hipMemcpyHostToDevice));
    return TOED_OK;
}

static int batch_enqueue(toed_batch *b, hipStream_t s) {
    ImageJob dummy;
    std::memset(&dummy, 0, sizeof dummy);
    return enqueue(s, b->in_type, b->tile_rows, b->d_jobs, b->n, dummy, b->groups.data(), (int)b->groups.size(),
                   b->ntiles, b->nchunks,
                   b->d_counts, b->d_offsets, b->d_chunks, StageBuf{b->d_srec, b->d_skey}, b->d_tilecnt,
                   b->persistent ? b->d_ticket : nullptr, nullptr, b->qad, b->g_counts, b->g_flags, b->ccap,
                   b->d_cbase);
}

int toed_batch_run(toed_batch *b, int *d_counts, int *d_flags) {
    if (!b) return TOED_ERR_ARG;
    HIPCHK(hipSetDevice(b->ctx->device));
    if (b->d_cbase && !d_counts) return TOED_ERR_ARG;    // compact output needs the per-image counts
    if (!b->uploaded || b->g_counts != d_counts || b->g_flags != d_flags || b->dirty) {
        b->dirty = false;
        HIPCHK(hipStreamSynchronize(b->ctx->stream));
        int st = batch_upload(b, d_counts, d_flags);
        if (st) return st;
        b->uploaded = true;
        if (b->gexec) { (void)hipGraphExecDestroy(b->gexec); b->gexec = nullptr; }
        if (b->graph) { (void)hipGraphDestroy(b->graph); b->graph = nullptr; }
        b->g_counts = d_counts; b->g_flags = d_flags;
    }
    return batch_enqueue(b, b->ctx->stream);
}

int toed_batch_graph_run(toed_batch *b, int *d_counts, int *d_flags) {
    if (!b) return TOED_ERR_ARG;
    HIPCHK(hipSetDevice(b->ctx->device));
    // the legacy default stream cannot be captured: capture and replay on the
    // context's own (blocking) stream, which orders with the default stream
    hipStream_t s = b->ctx->stream ? b->ctx->stream : b->ctx->own_stream;
    if (b->d_cbase && !d_counts) return TOED_ERR_ARG;
    if (!b->gexec || !b->uploaded || b->g_counts != d_counts || b->g_flags != d_flags || b->dirty) {
        b->dirty = false;
        HIPCHK(hipStreamSynchronize(s));
        int st = batch_upload(b, d_counts, d_flags);
        if (st) return st;
        b->uploaded = true;
        if (b->gexec) { (void)hipGraphExecDestroy(b->gexec); b->gexec = nullptr; }
        if (b->graph) { (void)hipGraphDestroy(b->graph); b->graph = nullptr; }
        b->g_counts = d_counts; b->g_flags = d_flags;     // the captured launches read these
        HIPCHK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        st = batch_enqueue(b, s);
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(s, &g);
        if (st) { if (g) (void)hipGraphDestroy(g); return st; }
        HIPCHK(e);
        b->graph = g;
        HIPCHK(hipGraphInstantiate(&b->gexec, g, nullptr, nullptr, 0));
    }
    HIPCHK(hipGraphLaunch(b->gexec, s));
    return TOED_OK;
}

// Times one configuration of a plan: a graph replay to capture and warm up,
// then reps replays between two events on the replay stream.  A run whose
// staging overflows (flags) is not a candidate: *ms = -1.
static int time_config(toed_batch *b, int reps, int *d_cf, hipEvent_t e0, hipEvent_t e1, float *ms) {
    *ms = -1.f;
    hipStream_t s = b->ctx->stream ? b->ctx->stream : b->ctx->own_stream;
    int st = toed_batch_graph_run(b, d_cf, d_cf + b->n);
    if (st) return st;
    std::vector<int> fl(b->n);
    HIPCHK(hipStreamSynchronize(s));
    HIPCHK(hipMemcpy(fl.data(), d_cf + b->n, sizeof(int) * b->n, hipMemcpyDeviceToHost));
    for (int f : fl)
        if (f & TOED_FLAG_STAGING) return TOED_OK;
    HIPCHK(hipEventRecord(e0, s));
    for (int r = 0; r < reps; ++r) {
        st = toed_batch_graph_run(b, d_cf, d_cf + b->n);
        if (st) return st;
    }
    HIPCHK(hipEventRecord(e1, s));
    HIPCHK(hipEventSynchronize(e1));
    float t = 0.f;
    HIPCHK(hipEventElapsedTime(&t, e0, e1));
    *ms = t / reps;
    return TOED_OK;
}

int toed_batch_autotune(toed_ctx *c, int n, const toed_image_t *imgs, int reps, toed_tune_t *best,
                        toed_tune_t *results, int max_results, int *n_results) {
    if (!c || n <= 0 || !imgs || !best || reps <= 0 || max_results < 0) return TOED_ERR_ARG;
    if (n_results) *n_results = 0;
    HIPCHK(hipSetDevice(c->device));
    int *d_cf = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (hipMalloc(&d_cf, sizeof(int) * 2 * n) != hipSuccess) return TOED_ERR_ALLOC;
    const int saved_rows = c->tile_rows;
    int st = (hipEventCreate(&e0) == hipSuccess && hipEventCreate(&e1) == hipSuccess) ? TOED_OK : TOED_ERR_HIP;
    toed_tune_t top{0, 0, 0, -1.f};
    int nres = 0;
    const int rows[2] = {8, 16};
    for (int k = 0; k < 2 && !st; ++k) {
        c->tile_rows = rows[k];
        toed_batch *b = nullptr;
        st = toed_batch_plan(c, n, imgs, &b);
        if (st) break;
        int occ = 1;
        for (const TileGroup &g : b->groups) occ = std::max(occ, g.occ);
        // one workgroup per tile, then persistent grids of 1 .. occ workgroups per CU
        for (int w = -1; w <= occ && !st; ++w) {
            if (w == 0) continue;
            st = toed_batch_set_persistent(b, w > 0);
            if (!st) st = toed_batch_set_grid(b, w > 0 ? w : 0);
            float ms = -1.f;
            if (!st) st = time_config(b, reps, d_cf, e0, e1, &ms);
            if (st || ms < 0.f) continue;
            const toed_tune_t r{rows[k], w > 0 ? 1 : 0, w > 0 ? w : 0, ms};
            if (results && nres < max_results) results[nres] = r;
            ++nres;
            if (top.ms < 0.f || ms < top.ms) top = r;
        }
        (void)toed_batch_destroy(b);
    }
    c->tile_rows = (!st && top.ms >= 0.f) ? top.tile_rows : saved_rows;
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipFree(d_cf);
    if (st) return st;
    if (top.ms < 0.f) return TOED_ERR_STAGING;   // every configuration overflowed its staging
    *best = top;
    if (n_results) *n_results = std::min(nres, max_results);
    return TOED_OK;
}

int toed_batch_backward(toed_batch *b, const toed_grad_out_t *outs) {
    if (!b || !outs) return TOED_ERR_ARG;
    HIPCHK(hipSetDevice(b->ctx->device));
    hipStream_t s = b->ctx->stream;
    std::vector<GradOut> h(b->n);
    for (int k = 0; k < b->n; ++k) {
        const toed_grad_out_t &o = outs[b->order[k]];
        if (!o.d_gout || !o.d_gimg || o.gpitch < b->jobs[k].W) return TOED_ERR_ARG;
        h[k].gout = o.d_gout; h[k].gimg = o.d_gimg; h[k].gsigma = o.d_gsigma; h[k].gpitch = o.gpitch;
    }
    if (!b->d_gouts) {
        if (hipMalloc(&b->d_gouts, sizeof(GradOut) * b->n) != hipSuccess) { b->d_gouts = nullptr; return TOED_ERR_ALLOC; }
        if (hipHostMalloc(&b->h_gouts, sizeof(GradOut) * b->n) != hipSuccess) { b->h_gouts = nullptr; return TOED_ERR_ALLOC; }
        HIPCHK(hipEventCreateWithFlags(&b->gouts_ev, hipEventDisableTiming));
        b->last_gouts.clear();
    }
    // The table is uploaded only when it changed (callers that reuse their
    // gradient buffers launch without any host wait); the pinned staging is
    // rewritten only after its previous copy has completed.
    if (b->last_gouts.size() != h.size() || std::memcmp(b->last_gouts.data(), h.data(), sizeof(GradOut) * b->n)) {
        HIPCHK(hipEventSynchronize(b->gouts_ev));
        std::memcpy(b->h_gouts, h.data(), sizeof(GradOut) * b->n);
        HIPCHK(hipMemcpyAsync(b->d_gouts, b->h_gouts, sizeof(GradOut) * b->n, hipMemcpyHostToDevice, s));
        HIPCHK(hipEventRecord(b->gouts_ev, s));
        b->last_gouts = h;
    }
    for (const TileGroup &g : b->groups) {
        if (g.tile_end <= g.tile_begin) continue;
        const dim3 grid((unsigned)(g.tile_end - g.tile_begin));
        ImageJob dummy;
        std::memset(&dummy, 0, sizeof dummy);
        for_tile_rows(b->tile_rows, [&](auto THc) {
            constexpr int T = decltype(THc)::value;
            for_radius(g.R1, [&](auto Rc) {
                constexpr int R = decltype(Rc)::value;
                if (b->in_type == TOED_U8)
                    hipLaunchKernelGGL((toed_tile_backward_kernel<uint8_t, T, R>), grid, dim3(NT), 0, s, b->d_jobs,
                                       b->n, dummy, b->d_gouts, StageBuf{b->d_srec, b->d_skey}, b->d_tilecnt,
                                       b->d_counts, b->d_offsets, b->d_chunks, g.tile_begin);
                else
                    hipLaunchKernelGGL((toed_tile_backward_kernel<float, T, R>), grid, dim3(NT), 0, s, b->d_jobs,
                                       b->n, dummy, b->d_gouts, StageBuf{b->d_srec, b->d_skey}, b->d_tilecnt,
                                       b->d_counts, b->d_offsets, b->d_chunks, g.tile_begin);
            });
        });
    }
    HIPCHK(hipGetLastError());
    return TOED_OK;
}

int toed_batch_destroy(toed_batch *b) {
    if (!b) return TOED_ERR_ARG;
    (void)hipSetDevice(b->ctx->device);
    (void)hipStreamSynchronize(b->ctx->stream);
    if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    (void)hipFree(b->d_jobs); (void)hipFree(b->d_counts); (void)hipFree(b->d_offsets); (void)hipFree(b->d_chunks);
    (void)hipFree(b->d_tilecnt); (void)hipFree(b->d_ticket); (void)hipFree(b->d_srec); (void)hipFree(b->d_skey);
    (void)hipFree(b->d_gouts);
    if (b->h_gouts) (void)hipHostFree(b->h_gouts);
    if (b->gouts_ev) (void)hipEventDestroy(b->gouts_ev);
    (void)hipFree(b->d_cbase);
    delete b;
    return TOED_OK;
}

// ---- unfused launch layer --------------------------------------------------
static TapSet *compat_taps(float sigma, int *st) {
    static std::vector<std::pair<float, TapSet *>> cache;
    std::lock_guard<std::mutex> lk(g_mu);
    int dev = 0;
    (void)hipGetDevice(&dev);
    for (auto &kv : cache) if (kv.first == sigma) return kv.second;   // single-device use
    int R0, R1;
    if ((*st = radius_for(sigma, &R0, &R1))) return nullptr;
    TapSet h;
    fill_taps(sigma, R0, R1, &h);
    TapSet *d = nullptr;
    if (hipMalloc(&d, sizeof(TapSet)) != hipSuccess) { *st = TOED_ERR_ALLOC; return nullptr; }
    if (hipMemcpy(d, &h, sizeof h, hipMemcpyHostToDevice) != hipSuccess) { *st = TOED_ERR_HIP; return nullptr; }
    cache.emplace_back(sigma, d);
    return d;
}

int toed_gpu_convolve_sigma_f32(int device_id, int h, int w, int ih, int iw, const float *img, float *Ix, float *Iy,
                                float *mag, float *orient, float sigma) {
    if (h <= 0 || w <= 0 || ih != 2 * h || iw != 2 * w || !img || !Ix || !Iy || !mag || !orient) return TOED_ERR_ARG;
    if ((long long)h * w > (1LL << 28)) return TOED_ERR_ARG;
    HIPCHK(hipSetDevice(device_id));
    int st = 0;
    TapSet *tp = compat_taps(sigma, &st);
    if (!tp) return st;
    const long long n = 4LL * h * w;
    hipLaunchKernelGGL(toed_planes_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, nullptr, h, w, img, Ix,
                       Iy, mag, orient, tp);
    HIPCHK(hipGetLastError());
    return TOED_OK;
}

int toed_gpu_nms_thresh_f32(int device_id, int ih, int iw, const float *Ix, const float *Iy, const float *mag, float *px,
                            float *py, float thresh, int border) {
    if (ih <= 0 || iw <= 0 || !Ix || !Iy || !mag || !px || !py || border < 1) return TOED_ERR_ARG;
    if ((long long)ih * iw > (1LL << 30)) return TOED_ERR_ARG;
    HIPCHK(hipSetDevice(device_id));
    const long long n = (long long)ih * iw;
    hipLaunchKernelGGL(toed_nms_maps_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, nullptr, ih, iw, Ix,
                       Iy, mag, px, py, thresh, border);
    HIPCHK(hipGetLastError());
    return TOED_OK;
}

int toed_gpu_convolve_sigma_f64(int device_id, int h, int w, int ih, int iw, const double *img, double *Ix,
                                double *Iy, double *mag, double *orient, double sigma) {
    if (h <= 0 || w <= 0 || ih != 2 * h || iw != 2 * w || !img || !Ix || !Iy || !mag || !orient) return TOED_ERR_ARG;
    if ((long long)h * w > (1LL << 28)) return TOED_ERR_ARG;
    HIPCHK(hipSetDevice(device_id));
    int st = 0;
    TapSet *tp = compat_taps((float)sigma, &st);
    if (!tp) return st;
    const long long n = 4LL * h * w;
    hipLaunchKernelGGL(toed_planes_f64_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, nullptr, h, w, img,
                       Ix, Iy, mag, orient, tp);
    HIPCHK(hipGetLastError());
    return TOED_OK;
}

int toed_gpu_nms_thresh_f64(int device_id, int ih, int iw, const double *Ix, const double *Iy, const double *mag,
                            double *px, double *py, double thresh, int border) {
    if (ih <= 0 || iw <= 0 || !Ix || !Iy || !mag || !px || !py || border < 1) return TOED_ERR_ARG;
    if ((long long)ih * iw > (1LL << 30)) return TOED_ERR_ARG;
    HIPCHK(hipSetDevice(device_id));
    const long long n = (long long)ih * iw;
    hipLaunchKernelGGL(toed_nms_maps_f64_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, nullptr, ih, iw,
                       Ix, Iy, mag, px, py, thresh, border);
    HIPCHK(hipGetLastError());
    return TOED_OK;
}


// The reference's own argument lists (gpu_kernels.hpp:24-63): eight device
// tap arrays for its fixed window (R0 = 8, R1 = 9), NMS with its literal
// threshold 2 and border 10 (gpu_nms.cu:97,208).  The taps go through one
// TapSet per device, packed on the null stream before the launch.
}  // extern "C"

static TapSet *g_ref_taps[NOALLOC_MAX_DEVICES];

static int ref_taps(int device_id, TapSet **out) {
    if (device_id < 0 || device_id >= NOALLOC_MAX_DEVICES) return TOED_ERR_ARG;
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_ref_taps[device_id]) {
        TapSet *d = nullptr;
        if (hipMalloc(&d, sizeof(TapSet)) != hipSuccess) return TOED_ERR_ALLOC;
        g_ref_taps[device_id] = d;
    }
    *out = g_ref_taps[device_id];
    return TOED_OK;
}

template <typename T>
static int convolve_taps(int device_id, int h, int w, int ih, int iw, const T *img, T *Ix, T *Iy, T *mag, T *orient,
                         const T *Gx, const T *Gxx, const T *Gxxx, const T *G, const T *Gx_sh, const T *Gxx_sh,
                         const T *Gxxx_sh, const T *G_sh) {
    if (h <= 0 || w <= 0 || ih != 2 * h || iw != 2 * w || !img || !Ix || !Iy || !mag || !orient) return TOED_ERR_ARG;
    if (!Gx || !Gxx || !Gxxx || !G || !Gx_sh || !Gxx_sh || !Gxxx_sh || !G_sh) return TOED_ERR_ARG;
    if ((long long)h * w > (1LL << 28)) return TOED_ERR_ARG;
    HIPCHK(hipSetDevice(device_id));
    TapSet *tp = nullptr;
    int st = ref_taps(device_id, &tp);
    if (st) return st;
    hipLaunchKernelGGL(toed_pack_taps_kernel<T>, dim3(1), dim3(64), 0, nullptr, tp, Gx, Gxx, Gxxx, G, Gx_sh, Gxx_sh,
                       Gxxx_sh, G_sh);
    const long long n = 4LL * h * w;
    if constexpr (std::is_same<T, float>::value)
        hipLaunchKernelGGL(toed_planes_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, nullptr, h, w, img, Ix,
                           Iy, mag, orient, tp);
    else
        hipLaunchKernelGGL(toed_planes_f64_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, nullptr, h, w, img,
                           Ix, Iy, mag, orient, tp);
    HIPCHK(hipGetLastError());
    return TOED_OK;
}

extern "C" {

int toed_gpu_convolve_f32(int device_id, int h, int w, int interp_h, int interp_w, const float *dev_img,
                          float *dev_Ix, float *dev_Iy, float *dev_I_grad_mag, float *dev_I_orient,
                          const float *dev_Gx, const float *dev_Gxx, const float *dev_Gxxx, const float *dev_G_of_x,
                          const float *dev_Gx_sh, const float *dev_Gxx_sh, const float *dev_Gxxx_sh,
                          const float *dev_G_of_x_sh) {
    return convolve_taps<float>(device_id, h, w, interp_h, interp_w, dev_img, dev_Ix, dev_Iy, dev_I_grad_mag,
                                dev_I_orient, dev_Gx, dev_Gxx, dev_Gxxx, dev_G_of_x, dev_Gx_sh, dev_Gxx_sh,
                                dev_Gxxx_sh, dev_G_of_x_sh);
}

int toed_gpu_convolve_f64(int device_id, int h, int w, int interp_h, int interp_w, const double *dev_img,
                          double *dev_Ix, double *dev_Iy, double *dev_I_grad_mag, double *dev_I_orient,
                          const double *dev_Gx, const double *dev_Gxx, const double *dev_Gxxx,
                          const double *dev_G_of_x, const double *dev_Gx_sh, const double *dev_Gxx_sh,
                          const double *dev_Gxxx_sh, const double *dev_G_of_x_sh) {
    return convolve_taps<double>(device_id, h, w, interp_h, interp_w, dev_img, dev_Ix, dev_Iy, dev_I_grad_mag,
                                 dev_I_orient, dev_Gx, dev_Gxx, dev_Gxxx, dev_G_of_x, dev_Gx_sh, dev_Gxx_sh,
                                 dev_Gxxx_sh, dev_G_of_x_sh);
}

int toed_gpu_nms_f32(int device_id, int interp_h, int interp_w, const float *dev_Ix, const float *dev_Iy,
                     const float *dev_I_grad_mag, float *dev_subpix_pos_x_map, float *dev_subpix_pos_y_map) {
    return toed_gpu_nms_thresh_f32(device_id, interp_h, interp_w, dev_Ix, dev_Iy, dev_I_grad_mag,
                                   dev_subpix_pos_x_map, dev_subpix_pos_y_map, 2.0f, 10);
}

int toed_gpu_nms_f64(int device_id, int interp_h, int interp_w, const double *dev_Ix, const double *dev_Iy,
                     const double *dev_I_grad_mag, double *dev_subpix_pos_x_map, double *dev_subpix_pos_y_map) {
    return toed_gpu_nms_thresh_f64(device_id, interp_h, interp_w, dev_Ix, dev_Iy, dev_I_grad_mag,
                                   dev_subpix_pos_x_map, dev_subpix_pos_y_map, 2.0, 10);
}

}  // extern "C"
