// TEST INFRASTRUCTURE ONLY -- a command-line harness around the reference's
// own SubpixelTOCorrectionCPU<double>, compiled from the reference sources in
// place by oracle/Makefile into oracle/_ref/ (never committed).  Used to write
// the golden vectors under tests/golden/ and, optionally, as the timed CPU
// baseline of the subpixel correction path.
//
// Drives the class as main_subpix_to_correction.cpp:279-289 does: ctor(H, W,
// thresh, sigma, nthreads, n), set_neighbor_radius, set_edgemap(double*),
// set_orientation(double*) or estimate_orientation(), run(edginfo) with a
// caller buffer of one row per work-grid pixel (:271 sizes it H*W x 4, enough
// for the driver's n = 0; run() writes up to work_height*work_width rows).
//
// usage: subpix_ref <map.f64> <orient.f64|-> <H> <W> <thresh> <sigma>
//                   <nthreads> <nbr_radius> <out.bin> <tmp_dir> [interp_n]
// out.bin: int64 N, double t_nms, t_conv, t_correct, t_total (s), N x 4
// doubles, then TO_edgemap and TO_orientation (H x W doubles each).
#include <omp.h>

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "subpix_to_correction.hpp"

static bool read_f64(const char *path, std::vector<double> &v, size_t n) {
    std::FILE *f = std::fopen(path, "rb");
    if (!f) return false;
    v.resize(n);
    const size_t got = std::fread(v.data(), sizeof(double), n, f);
    std::fclose(f);
    return got == n;
}

int main(int argc, char **argv) {
    if (argc != 11 && argc != 12) {
        std::fprintf(stderr, "usage: %s map orient|- H W thresh sigma nthreads nbr_radius out.bin tmpdir [interp_n]\n",
                     argv[0]);
        return 2;
    }
    const int interp_n = argc == 12 ? std::atoi(argv[11]) : 0;
    if (interp_n < 0 || interp_n > 8) return 2;
    const int H = std::atoi(argv[3]), W = std::atoi(argv[4]);
    const double thresh = std::atof(argv[5]), sigma = std::atof(argv[6]);
    const int nt = std::atoi(argv[7]);
    const double nbr = std::atof(argv[8]);
    if (H <= 0 || W <= 0 || !(sigma > 0)) return 2;
    std::vector<double> E, O;
    if (!read_f64(argv[1], E, (size_t)H * W)) return 2;
    const bool have_o = std::strcmp(argv[2], "-") != 0;
    if (have_o && !read_f64(argv[2], O, (size_t)H * W)) return 2;

    const size_t work_px = (size_t)((H - 1) * (1 << interp_n) + 1) * (size_t)((W - 1) * (1 << interp_n) + 1);
    std::vector<double> edginfo(work_px * 4, 0.0);
    const double t0 = omp_get_wtime();
    SubpixelTOCorrectionCPU<double> corr(H, W, thresh, sigma, nt < 1 ? 1 : nt, interp_n);
    corr.set_output_dir(argv[10]);
    corr.set_neighbor_radius(nbr);
    corr.set_edgemap(E.data());
    if (have_o) corr.set_orientation(O.data());
    else corr.estimate_orientation();
    const int n = corr.run(edginfo.data());
    const double t1 = omp_get_wtime();

    std::FILE *f = std::fopen(argv[9], "wb");
    if (!f) return 3;
    const int64_t n64 = n;
    const double times[4] = {corr.time_nms, corr.time_conv, corr.time_correct, t1 - t0};
    std::fwrite(&n64, sizeof n64, 1, f);
    std::fwrite(times, sizeof(double), 4, f);
    std::fwrite(edginfo.data(), sizeof(double), (size_t)n * 4, f);
    std::fwrite(corr.TO_edgemap, sizeof(double), (size_t)H * W, f);
    std::fwrite(corr.TO_orientation, sizeof(double), (size_t)H * W, f);
    std::fclose(f);
    return 0;
}
