// toed_subpix_gpu.hpp -- header-only drop-in for the reference's
// SubpixelTOCorrectionCPU<T> (subpix_TO_correction/subpix_to_correction.hpp:20-96)
// running on the GPU through the C ABI of libtoed_hip.so (toed_capi.h).
// Same constructor argument order (H, W, threshold, g_sigma, cpu_nthreads, n),
// same methods (set_edgemap, set_orientation, estimate_orientation, run,
// set_output_dir, set_neighbor_radius, write_array_to_file) and public fields
// (subpix_edge_pts_final, TO_edgemap, TO_orientation, edge_pt_list_idx,
// num_of_edge_data, time_nms, time_conv, time_correct), so the reference's
// driver (main_subpix_to_correction.cpp:279-293) compiles against it with the
// class name changed.  Link with -ltoed_hip -lamdhip64.
//
// Differences from the reference, by design:
//   * The arithmetic is FP32 for the map, the planes and the convolutions,
//     FP64 for the orientation estimate, the NMS decision and parabola, the
//     cubic sampling and the orientation (T = double inputs are rounded to
//     float once when uploaded; outputs are converted back to T).
//   * cpu_nthreads is accepted and ignored; the GPU is the current HIP device.
//   * n = interp_n > 0 (the bilinear work grid of :92-105, :226-240) at most
//     TOED_SUBPIX_MAX_INTERP; the map is resized on the GPU from its float
//     upload (FP64 arithmetic, the work grid stored as float).
//   * time_nms / time_conv / time_correct are in ms (GPU events): the
//     orientation + NMS kernel, the isolation and convolution + correction
//     kernels (which write the list rows), and the rank scans + TO maps
//     (the reference: seconds of CPU time).
#ifndef TOED_SUBPIX_GPU_HPP
#define TOED_SUBPIX_GPU_HPP

#include <hip/hip_runtime_api.h>

#include <cstdio>
#include <fstream>
#include <iostream>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "toed_capi.h"

template <typename T>
class SubpixelTOCorrectionGPU {
    static_assert(std::is_same<T, float>::value || std::is_same<T, double>::value, "T = float or double");
    int img_height, img_width;
    int interp_n, work_height, work_width;
    T thresh, sigma, neighbor_radius;
    std::string output_dir;
    toed_ctx *ctx = nullptr;
    float *d_map = nullptr, *d_orient = nullptr, *d_edges = nullptr, *d_te = nullptr, *d_to = nullptr;
    int *d_count = nullptr;
    bool have_orient = false;
    std::vector<T> edges_host, te_host, to_host;

    static void check(int st, const char *what) {
        if (st != TOED_OK) {
            std::fprintf(stderr, "toed: %s failed with status %d\n", what, st);
            throw std::runtime_error(what);
        }
    }
    static void hcheck(hipError_t e, const char *what) {
        if (e != hipSuccess) check(TOED_ERR_HIP, what);
    }
    void upload(const T *data, float *dst) {
        std::vector<float> f((size_t)img_height * img_width);
        for (size_t k = 0; k < f.size(); ++k) f[k] = (float)data[k];
        hcheck(hipMemcpy(dst, f.data(), f.size() * sizeof(float), hipMemcpyHostToDevice), "upload");
    }

  public:
    T *subpix_edge_pts_final = nullptr;
    T *TO_edgemap = nullptr;
    T *TO_orientation = nullptr;
    int edge_pt_list_idx = 0;
    int num_of_edge_data = 4;
    double time_nms = 0, time_conv = 0, time_correct = 0;

    SubpixelTOCorrectionGPU(int H, int W, T threshold, T g_sigma, int cpu_nthreads, int n = 0)
        : img_height(H), img_width(W), interp_n(n < 0 ? 0 : n), thresh(threshold), sigma(g_sigma),
          neighbor_radius(2), output_dir("./output_files") {
        (void)cpu_nthreads;
        if (interp_n > TOED_SUBPIX_MAX_INTERP) check(TOED_ERR_ARG, "interp_n");
        work_height = (H - 1) * (1 << interp_n) + 1;
        work_width = (W - 1) * (1 << interp_n) + 1;
        int R = 0;
        check(toed_subpix_radius((double)g_sigma, &R), "sigma");
        int dev = 0;
        hcheck(hipGetDevice(&dev), "hipGetDevice");
        check(toed_create(dev, &ctx), "toed_create");
        const size_t px = (size_t)H * W;
        hcheck(hipMalloc(&d_map, px * sizeof(float)), "hipMalloc");
        hcheck(hipMalloc(&d_orient, px * sizeof(float)), "hipMalloc");
        const size_t wpx = (size_t)work_height * work_width;    // one list row per work pixel at most
        hcheck(hipMalloc(&d_edges, wpx * 4 * sizeof(float)), "hipMalloc");
        hcheck(hipMalloc(&d_te, px * sizeof(float)), "hipMalloc");
        hcheck(hipMalloc(&d_to, px * sizeof(float)), "hipMalloc");
        hcheck(hipMalloc(&d_count, sizeof(int)), "hipMalloc");
        hcheck(hipMemset(d_map, 0, px * sizeof(float)), "hipMemset");
        edges_host.assign(wpx * 4, T(0));
        te_host.assign(px, T(0));
        to_host.assign(px, T(0));
        subpix_edge_pts_final = edges_host.data();
        TO_edgemap = te_host.data();
        TO_orientation = to_host.data();
    }
    ~SubpixelTOCorrectionGPU() {
        hipFree(d_map);
        hipFree(d_orient);
        hipFree(d_edges);
        hipFree(d_te);
        hipFree(d_to);
        hipFree(d_count);
        if (ctx) toed_destroy(ctx);
    }
    SubpixelTOCorrectionGPU(const SubpixelTOCorrectionGPU &) = delete;
    SubpixelTOCorrectionGPU &operator=(const SubpixelTOCorrectionGPU &) = delete;

    void set_edgemap(const T *data) { upload(data, d_map); }
    void set_orientation(const T *data) {
        upload(data, d_orient);
        have_orient = true;
    }
    void estimate_orientation() { have_orient = false; }
    void set_output_dir(const std::string &dir) { output_dir = dir.empty() ? "./output_files" : dir; }
    void set_neighbor_radius(T radius) { neighbor_radius = radius; }

    int run(T *edginfo_out) {
        toed_subpix_map_t m{};
        m.d_map = d_map;
        m.d_orient = have_orient ? d_orient : nullptr;
        m.H = img_height;
        m.W = img_width;
        m.pitch = m.opitch = m.tpitch = img_width;
        m.thresh = (double)thresh;
        m.d_edges = d_edges;
        m.capacity = (long long)work_height * work_width;
        m.d_to_edgemap = d_te;
        m.d_to_orientation = d_to;
        toed_subpix *p = nullptr;
        check(toed_subpix_plan_n(ctx, 1, &m, (double)sigma, (double)neighbor_radius, interp_n, &p),
              "toed_subpix_plan_n");
        check(toed_subpix_set_timing(p, 1), "toed_subpix_set_timing");
        int st = toed_subpix_run(p, d_count, nullptr);
        float a = 0.f, b = 0.f, c = 0.f;
        if (st == TOED_OK) st = toed_subpix_last_times(p, &a, &b, &c);
        toed_subpix_destroy(p);
        check(st, "toed_subpix_run");
        int n = 0;
        hcheck(hipMemcpy(&n, d_count, sizeof(int), hipMemcpyDeviceToHost), "count");
        const size_t px = (size_t)img_height * img_width;
        std::vector<float> e((size_t)n * 4), te(px), to(px);
        if (n) hcheck(hipMemcpy(e.data(), d_edges, e.size() * sizeof(float), hipMemcpyDeviceToHost), "edges");
        hcheck(hipMemcpy(te.data(), d_te, px * sizeof(float), hipMemcpyDeviceToHost), "TO_edgemap");
        hcheck(hipMemcpy(to.data(), d_to, px * sizeof(float), hipMemcpyDeviceToHost), "TO_orientation");
        for (size_t k = 0; k < e.size(); ++k) edges_host[k] = (T)e[k];
        for (size_t k = 0; k < px; ++k) {
            te_host[k] = (T)te[k];
            to_host[k] = (T)to[k];
        }
        if (edginfo_out)
            for (size_t k = 0; k < e.size(); ++k) edginfo_out[k] = (T)e[k];
        edge_pt_list_idx = n;
        time_nms = a;
        time_conv = b;
        time_correct = c;
        return n;
    }

    // tab-separated text, as subpix_to_correction.cpp:741-775
    void write_array_to_file(std::string filename, T *wr_data, int first_dim, int second_dim) {
        if (wr_data == nullptr || first_dim <= 0 || second_dim <= 0) return;
        std::string path = output_dir;
        if (!path.empty() && path.back() != '/') path.push_back('/');
        path += filename;
        std::ofstream out(path);
        if (!out.is_open()) return;
        for (int i = 0; i < first_dim; ++i) {
            for (int j = 0; j < second_dim; ++j) out << wr_data[(size_t)i * second_dim + j] << "\t";
            out << "\n";
        }
    }
};

#endif
