// Drives the header-only ThirdOrderEdgeDetectionGPU<T> (include/toed_gpu.hpp)
// the way the reference's main_gpu_cpu.cpp:195-209 drives its GPU class, once
// with T = float and once with T = double, on a raw H x W u8 image:
//
//     toed_class_driver <raw image> <H> <W> <sigma> <out dir>
//
// For each T it writes the class's text list into <out dir>/f32 or /f64
// (those directories must exist) and the returned rows, as T, into
// <out dir>/edges_f32.bin / edges_f64.bin, and prints the edgel counts.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <string>
#include <vector>

#include "toed_gpu.hpp"

template <typename T>
static int run(const char *raw, int H, int W, int sigma, const std::string &out, const char *tag) {
    std::ifstream infile(raw, std::ios::binary);
    if (!infile) {
        std::printf("cannot open %s\n", raw);
        return -1;
    }
    const int kernel_size = 2 * (int)std::ceil(4.0 * sigma) + 1;
    ThirdOrderEdgeDetectionGPU<T> toed(0, H, W, kernel_size, sigma);
    toed.set_output_dir(out + "/" + tag);
    toed.preprocessing(infile);
    toed.convolve_img();
    std::vector<T> edges((size_t)4 * (2 * H) * (2 * W));
    const int n = toed.non_maximum_suppresion(edges.data());
    std::printf("%s edges = %d\n", tag, n);
    std::ofstream bin(out + "/edges_" + tag + ".bin", std::ios::binary);
    bin.write((const char *)edges.data(), (std::streamsize)(sizeof(T) * 4 * (size_t)n));
    return n;
}

int main(int argc, char **argv) {
    if (argc != 6) {
        std::printf("usage: %s <raw u8 image> <H> <W> <sigma> <out dir>\n", argv[0]);
        return 2;
    }
    const int H = std::atoi(argv[2]), W = std::atoi(argv[3]), sigma = std::atoi(argv[4]);
    const std::string out = argv[5];
    if (run<float>(argv[1], H, W, sigma, out, "f32") < 0) return 1;
    if (run<double>(argv[1], H, W, sigma, out, "f64") < 0) return 1;
    return 0;
}
