// Drives the header-only SubpixelTOCorrectionGPU<T> (include/toed_subpix_gpu.hpp)
// the way the reference's main_subpix_to_correction.cpp:279-293 drives its
// SubpixelTOCorrectionCPU<double>: ctor(H, W, thresh, sigma, nthreads, n),
// set_output_dir, set_neighbor_radius, set_edgemap, set_orientation or
// estimate_orientation, run(edginfo) with a caller buffer of one row per
// work-grid pixel (H*W x 4 rows when n = 0, as :271), then
// write_array_to_file of the list.
//
//     toed_subpix_driver <map.f64> <orient.f64|-> <H> <W> <thresh> <sigma>
//                        <nbr_radius> <float|double> <out.bin> <out_dir> [interp_n]
//
// out.bin: int64 N, N x 4 doubles (the returned rows, as T, widened), then
// TO_edgemap and TO_orientation (H x W doubles each).
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "toed_subpix_gpu.hpp"

static bool read_f64(const char *path, std::vector<double> &v, size_t n) {
    std::FILE *f = std::fopen(path, "rb");
    if (!f) return false;
    v.resize(n);
    const size_t got = std::fread(v.data(), sizeof(double), n, f);
    std::fclose(f);
    return got == n;
}

template <typename T>
static int run(const std::vector<double> &E, const std::vector<double> &O, bool have_o, int H, int W, double thresh,
               double sigma, double nbr, const char *out, const char *dir, int interp_n) {
    std::vector<T> e(E.begin(), E.end()), o(O.begin(), O.end());
    const size_t wpx = (size_t)((H - 1) * (1 << interp_n) + 1) * (size_t)((W - 1) * (1 << interp_n) + 1);
    std::vector<T> edginfo(wpx * 4, T(0));
    SubpixelTOCorrectionGPU<T> to_corr(H, W, (T)thresh, (T)sigma, 1, interp_n);
    to_corr.set_output_dir(dir);
    to_corr.set_neighbor_radius((T)nbr);
    to_corr.set_edgemap(e.data());
    if (have_o) to_corr.set_orientation(o.data());
    else to_corr.estimate_orientation();
    const int n = to_corr.run(edginfo.data());
    std::printf("Number of corrected edgels = %d\n", n);
    to_corr.write_array_to_file("TO_corrected_edges.txt", edginfo.data(), n, 4);
    std::FILE *f = std::fopen(out, "wb");
    if (!f) return 3;
    const int64_t n64 = n;
    std::fwrite(&n64, sizeof n64, 1, f);
    std::vector<double> rows(edginfo.begin(), edginfo.begin() + (size_t)n * 4);
    std::vector<double> te(to_corr.TO_edgemap, to_corr.TO_edgemap + (size_t)H * W);
    std::vector<double> to(to_corr.TO_orientation, to_corr.TO_orientation + (size_t)H * W);
    std::fwrite(rows.data(), sizeof(double), rows.size(), f);
    std::fwrite(te.data(), sizeof(double), te.size(), f);
    std::fwrite(to.data(), sizeof(double), to.size(), f);
    std::fclose(f);
    return 0;
}

int main(int argc, char **argv) {
    if (argc != 11 && argc != 12) {
        std::fprintf(stderr,
                     "usage: %s map orient|- H W thresh sigma nbr_radius float|double out.bin out_dir [interp_n]\n",
                     argv[0]);
        return 2;
    }
    const int interp_n = argc == 12 ? std::atoi(argv[11]) : 0;
    if (interp_n < 0 || interp_n > TOED_SUBPIX_MAX_INTERP) return 2;
    const int H = std::atoi(argv[3]), W = std::atoi(argv[4]);
    const double thresh = std::atof(argv[5]), sigma = std::atof(argv[6]), nbr = std::atof(argv[7]);
    if (H <= 0 || W <= 0 || !(sigma > 0)) return 2;
    std::vector<double> E, O;
    if (!read_f64(argv[1], E, (size_t)H * W)) return 2;
    const bool have_o = std::strcmp(argv[2], "-") != 0;
    if (have_o && !read_f64(argv[2], O, (size_t)H * W)) return 2;
    try {
        if (std::strcmp(argv[8], "float") == 0)
            return run<float>(E, O, have_o, H, W, thresh, sigma, nbr, argv[9], argv[10], interp_n);
        return run<double>(E, O, have_o, H, W, thresh, sigma, nbr, argv[9], argv[10], interp_n);
    } catch (const std::exception &ex) {
        std::fprintf(stderr, "toed_subpix_driver: %s\n", ex.what());
        return 4;
    }
}
