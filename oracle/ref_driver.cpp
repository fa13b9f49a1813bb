// TEST INFRASTRUCTURE ONLY -- a command-line harness around the reference's
// own FP64 CPU detector, compiled from the reference sources in place by
// oracle/Makefile into oracle/_ref/ (never committed).  Used to write the
// golden vectors under tests/golden/ and, optionally, as the timed CPU
// baseline.
//
// Drives ThirdOrderEdgeDetectionCPU<double> exactly as main_cpu_DP.cpp:196-204
// does: ctor(H, W, sigma = 2, kernel_size = 17, nthreads), preprocessing from
// a byte stream (cpu_toed.cpp:97-126, the reference's no-OpenCV input path),
// convolve_img(), non_maximum_suppresion(TOED_edges) with a caller buffer of
// (2H)(2W) x 4 rows (main_gpu_cpu.cpp:56-66).
//
// usage: toed_ref <raw_u8_file> <H> <W> <nthreads> <out.bin> <tmp_dir>
// out.bin: int64 N, double time_conv_s, double time_nms_s, then N x 4 doubles.
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <vector>

#include "cpu_toed.hpp"

int main(int argc, char **argv) {
    if (argc != 7) {
        std::fprintf(stderr, "usage: %s raw H W nthreads out.bin tmpdir\n", argv[0]);
        return 2;
    }
    const int H = std::atoi(argv[2]), W = std::atoi(argv[3]), nt = std::atoi(argv[4]);
    std::ifstream in(argv[1], std::ios::binary);
    if (!in || H <= 0 || W <= 0) {
        std::fprintf(stderr, "bad input\n");
        return 2;
    }
    ThirdOrderEdgeDetectionCPU<double> det(H, W, 2, 17, nt);
    det.set_output_dir(argv[6]);
    det.preprocessing(in);
    det.convolve_img();
    std::vector<double> edges((size_t)(2 * H) * (2 * W) * 4, 0.0);
    const int n = det.non_maximum_suppresion(edges.data());
    std::FILE *f = std::fopen(argv[5], "wb");
    if (!f) return 3;
    const int64_t n64 = n;
    const double tc = det.time_conv, tn = det.time_nms;
    std::fwrite(&n64, sizeof n64, 1, f);
    std::fwrite(&tc, sizeof tc, 1, f);
    std::fwrite(&tn, sizeof tn, 1, f);
    std::fwrite(edges.data(), sizeof(double), (size_t)n * 4, f);
    std::fclose(f);
    return 0;
}
