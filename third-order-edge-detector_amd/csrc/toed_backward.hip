// Differentiable TOED: adjoint of the detection outputs with the edgel set
// frozen (no reference counterpart; the forward arithmetic it differentiates
// is cpu_toed.cpp:194-232 and :411-489).
//
// For every edgel of a forward pass (its interp pixel (R, C) from the forward's
// index output) one group of toedadj::G lanes of toed_backward_kernel:
//   1. recomputes in FP64 from the image the nine responses at the edgel and
//      Ix, Iy at its four interpolation neighbours (octant from the recomputed
//      gradient), each with its derivative with respect to sigma (window radii
//      frozen), lanes splitting the window rows, reduced with DPP group sums;
//   2. runs the hand-written reverse pass of the parabola fit, the subpixel
//      position, the strength and the third-order orientation for the upstream
//      row (dL/dx, dL/dy, dL/dtheta, dL/dstrength), giving 17 adjoints
//      (9 at the edgel, Ix and Iy at each neighbour);
//   3. adds sum(adjoint * d response / d sigma) to dL/dsigma (FP64 atomic);
//   4. scatters the transposed filters: lanes own rows of the (2R1+3)^2 patch
//      around the edgel, sum the five points' separable contributions in FP64
//      and add them to dL/dimage with hardware float atomics.
// Derivative planes are never stored: everything is recomputed from the image.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <vector>

#include "toed_capi.h"
#include "toed_adjoint.hpp"

namespace toedbw {

constexpr int NT = 256;
constexpr int G = toedadj::G;         // lanes per edgel
constexpr int MAXR1 = 9;
constexpr int MAXN = 2 * MAXR1 + 1;

// FP64 taps for one sigma and frozen radii: variant 0 unshifted (window R1),
// 1 unshifted with zero end taps (window R0, phase (0,0)), 2 shifted; families
// G, G', G'', G'''; and the same for d/dsigma.
struct BwTaps {
    int R0, R1, pad0, pad1;
    double t[3][4][MAXN];
    double d[3][4][MAXN];
};

struct BwJob {
    const void *img;
    const int *index;
    const int *count;
    const float *gout;
    float *gimg;
    double *gsigma;
    const BwTaps *taps;
    int in_type, H, W, pitch;
    int gpitch, R0, R1, max_edgels;
    int wg_begin, pad0, pad1, pad2;
};

#define TOED_CONST __attribute__((address_space(4)))

__device__ __forceinline__ BwJob load_bwjob(const BwJob *jobs, int njobs, int wg) {
#if defined(__HIP_DEVICE_COMPILE__)
    const TOED_CONST BwJob *cj = (const TOED_CONST BwJob *)jobs;
    int lo = 0, hi = njobs - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (cj[mid].wg_begin <= wg) lo = mid; else hi = mid - 1;
    }
    return cj[lo];
#else
    (void)njobs; (void)wg;
    return jobs[0];
#endif
}

template <typename IN_T> __device__ __forceinline__ double pix(const BwJob &J, int i, int j) {
    if (i < 0 || i >= J.H || j < 0 || j >= J.W) return 0.0;
    return (double)((const IN_T *)J.img)[(long long)i * J.pitch + j];
}

template <typename IN_T> struct GlobalImg {
    const BwJob *J;
    __device__ double operator()(int i, int j) const { return pix<IN_T>(*J, i, j); }
};

struct GlobalSink {
    float *g;
    int pitch;
    __device__ void operator()(int i, int j, float v) const { unsafeAtomicAdd(g + (long long)i * pitch + j, v); }
};

// One edgel of job J (lane group, lane l): see toed_adjoint.hpp.
template <typename IN_T, int R1>
__device__ __forceinline__ void edgel_adjoint(const BwJob &J, int e, int l, const double (*stp)[2 * R1 + 3][4],
                                              const double (*sdp)[2 * R1 + 3][4], const float4 (*stf4)[2 * R1 + 3],
                                              double *gs_out) {
    const int IW = 2 * J.W;
    const int idx = J.index[e];
    const int R = idx / IW, C = idx - (idx / IW) * IW;
    const float4 up = reinterpret_cast<const float4 *>(J.gout)[e];
    GlobalImg<IN_T> img{&J};
    GlobalSink sink{J.gimg, J.gpitch};
    toedadj::edgel_backward<R1>(R, C, up, l, J.H, J.W, stp, sdp, stf4, img, sink, gs_out);
}

template <typename IN_T, int R1>
__global__ __launch_bounds__(NT) void toed_backward_kernel(const BwJob *jobs, int njobs) {
    constexpr int N = 2 * R1 + 1;
    __shared__ double stp[3][N + 2][4];   // taps, zero border
    __shared__ double sdp[3][N + 2][4];   // d taps / d sigma
    __shared__ float4 stf4[3][N + 2];     // FP32 taps for the scatter
    __shared__ double sgs[NT / G];     // per-group dL/dsigma, reduced once per workgroup
    const int wg = blockIdx.x;
    const BwJob J = load_bwjob(jobs, njobs, wg);
    toedadj::tap_tables_fill<R1>(stp, sdp, stf4, threadIdx.x, NT,
                                 [&](int v, int f, int t) { return J.taps->t[v][f][t]; },
                                 [&](int v, int f, int t) { return J.taps->d[v][f][t]; });
    __syncthreads();
    const int n = min(*J.count, J.max_edgels);
    const int g = threadIdx.x / G, l = threadIdx.x % G;
    const int e = (wg - J.wg_begin) * (NT / G) + g;
    if (l == 0) sgs[g] = 0.0;
    if (e < n) edgel_adjoint<IN_T, R1>(J, e, l, stp, sdp, stf4, &sgs[g]);
    __syncthreads();
    if (threadIdx.x == 0 && J.gsigma) {
        double t = 0.0;
        for (int k = 0; k < NT / G; ++k) t += sgs[k];
        if (t != 0.0) unsafeAtomicAdd(J.gsigma, t);
    }
}

}  // namespace toedbw

using namespace toedbw;

namespace {

int radius_of(float sigma, int *R0, int *R1) {
    if (!(sigma > 0.f) || !std::isfinite(sigma)) return TOED_ERR_SIGMA;
    const int r0 = (int)std::ceil(4.0 * (double)sigma - 1e-9);
    if (r0 + 1 != 5 && r0 + 1 != 7 && r0 + 1 != 9) return TOED_ERR_SIGMA;
    *R0 = r0;
    *R1 = r0 + 1;
    return TOED_OK;
}

void fill_bw_taps(double sigma, int R0, int R1, BwTaps *T) {
    std::memset(T, 0, sizeof(*T));
    T->R0 = R0;
    T->R1 = R1;
    const int n = 2 * R1 + 1;
    const double s = sigma, s2 = s * s, norm = std::sqrt(2.0 * M_PI) * s;
    for (int sh = 0; sh < 2; ++sh) {
        for (int k = 0; k < n; ++k) {
            const double x = (double)(k - R1) + (sh ? 0.5 : 0.0);
            const double g = std::exp(-x * x / (2.0 * s2)) / norm;
            const double P[4] = {1.0, -x / s2, (x * x - s2) / (s2 * s2), x * (3.0 * s2 - x * x) / (s2 * s2 * s2)};
            const double dP[4] = {0.0, 2.0 * x / (s2 * s), -4.0 * x * x / (s2 * s2 * s) + 2.0 / (s2 * s),
                                  -12.0 * x / (s2 * s2 * s) + 6.0 * x * x * x / (s2 * s2 * s2 * s)};
            const double dg = g * (x * x / (s2 * s) - 1.0 / s);
            for (int f = 0; f < 4; ++f) {
                const double v = P[f] * g, dv = dP[f] * g + P[f] * dg;
                if (sh) {
                    T->t[2][f][k] = v; T->d[2][f][k] = dv;
                } else {
                    T->t[0][f][k] = v; T->d[0][f][k] = dv;
                    const bool end = (k == 0 || k == n - 1);
                    T->t[1][f][k] = end ? 0.0 : v; T->d[1][f][k] = end ? 0.0 : dv;
                }
            }
        }
    }
}

struct BwTapCache {
    std::mutex mu;
    std::vector<std::pair<std::pair<int, float>, BwTaps *>> items;   // (device, sigma) -> taps
};
BwTapCache g_bw_taps;

int bw_taps_for(int device, float sigma, const BwTaps **out) {
    std::lock_guard<std::mutex> lk(g_bw_taps.mu);
    for (auto &kv : g_bw_taps.items)
        if (kv.first.first == device && kv.first.second == sigma) { *out = kv.second; return TOED_OK; }
    int R0, R1;
    int st = radius_of(sigma, &R0, &R1);
    if (st) return st;
    BwTaps h;
    fill_bw_taps(sigma, R0, R1, &h);
    BwTaps *d = nullptr;
    if (hipMalloc(&d, sizeof(BwTaps)) != hipSuccess) return TOED_ERR_ALLOC;
    if (hipMemcpy(d, &h, sizeof h, hipMemcpyHostToDevice) != hipSuccess) { (void)hipFree(d); return TOED_ERR_HIP; }
    g_bw_taps.items.push_back({{device, sigma}, d});
    *out = d;
    return TOED_OK;
}

}  // namespace

extern "C" int toed_backward(toed_ctx *ctx, int n, const toed_grad_t *items) {
    if (!ctx || n <= 0 || !items) return TOED_ERR_ARG;
    const int device = toed_get_device(ctx);
    if (hipSetDevice(device) != hipSuccess) return TOED_ERR_HIP;
    hipStream_t s = (hipStream_t)toed_get_stream(ctx);
    // jobs grouped by (input type, radius): one specialised launch per group
    std::vector<BwJob> jobs;
    std::vector<int> r1s, types;
    for (int m = 0; m < n; ++m) {
        const toed_grad_t &it = items[m];
        if (it.max_edgels < 0) return TOED_ERR_ARG;
        if (it.max_edgels == 0) continue;          // nothing to back-propagate
        if (!it.d_img || !it.d_index || !it.d_count || !it.d_gout || !it.d_gimg || it.H <= 0 || it.W <= 0 ||
            it.pitch < it.W || it.gpitch < it.W || (it.in_type != TOED_U8 && it.in_type != TOED_F32))
            return TOED_ERR_ARG;
        BwJob J;
        std::memset(&J, 0, sizeof J);
        int st = radius_of(it.sigma, &J.R0, &J.R1);
        if (st) return st;
        const BwTaps *tp = nullptr;
        if ((st = bw_taps_for(device, it.sigma, &tp))) return st;
        J.img = it.d_img; J.index = it.d_index; J.count = it.d_count; J.gout = it.d_gout;
        J.gimg = it.d_gimg; J.gsigma = it.d_gsigma; J.taps = tp;
        J.in_type = it.in_type; J.H = it.H; J.W = it.W; J.pitch = it.pitch; J.gpitch = it.gpitch;
        J.max_edgels = it.max_edgels;
        jobs.push_back(J);
    }
    n = (int)jobs.size();
    if (n == 0) return TOED_OK;
    std::vector<int> order(n);
    for (int m = 0; m < n; ++m) order[m] = m;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
        return std::make_pair(jobs[a].in_type, jobs[a].R1) < std::make_pair(jobs[b].in_type, jobs[b].R1);
    });
    std::vector<BwJob> sorted;
    for (int m : order) sorted.push_back(jobs[m]);
    // one upload of the job table (pageable copy: completes before the launches use it)
    BwJob *d_jobs = nullptr;
    if (hipMallocAsync((void **)&d_jobs, sizeof(BwJob) * n, s) != hipSuccess) return TOED_ERR_ALLOC;
    size_t k = 0;
    std::vector<std::pair<size_t, size_t>> groups;
    while (k < sorted.size()) {
        size_t e = k;
        int wg = 0;
        while (e < sorted.size() && sorted[e].in_type == sorted[k].in_type && sorted[e].R1 == sorted[k].R1) {
            sorted[e].wg_begin = wg;
            wg += (sorted[e].max_edgels + (NT / G) - 1) / (NT / G);
            ++e;
        }
        groups.push_back({k, e});
        k = e;
    }
    if (hipMemcpyAsync(d_jobs, sorted.data(), sizeof(BwJob) * n, hipMemcpyHostToDevice, s) != hipSuccess)
        return TOED_ERR_HIP;
    for (auto &gr : groups) {
        const BwJob &J0 = sorted[gr.first];
        const BwJob &JL = sorted[gr.second - 1];
        const int nwg = JL.wg_begin + (JL.max_edgels + (NT / G) - 1) / (NT / G);
        if (nwg <= 0) continue;
        const BwJob *dj = d_jobs + gr.first;
        const int nj = (int)(gr.second - gr.first);
        auto launch = [&](auto kern) { hipLaunchKernelGGL(kern, dim3(nwg), dim3(NT), 0, s, dj, nj); };
        if (J0.in_type == TOED_U8) {
            if (J0.R1 == 5) launch(toed_backward_kernel<uint8_t, 5>);
            else if (J0.R1 == 7) launch(toed_backward_kernel<uint8_t, 7>);
            else launch(toed_backward_kernel<uint8_t, 9>);
        } else {
            if (J0.R1 == 5) launch(toed_backward_kernel<float, 5>);
            else if (J0.R1 == 7) launch(toed_backward_kernel<float, 7>);
            else launch(toed_backward_kernel<float, 9>);
        }
    }
    if (hipFreeAsync(d_jobs, s) != hipSuccess) return TOED_ERR_HIP;
    return hipGetLastError() == hipSuccess ? TOED_OK : TOED_ERR_HIP;
}
