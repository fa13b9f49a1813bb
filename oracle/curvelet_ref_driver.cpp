// TEST INFRASTRUCTURE ONLY -- a command-line harness around the reference's
// header-only curvelet code (curvelet/*.hpp), compiled from the reference
// sources in place by oracle/Makefile into oracle/_ref/ (never committed).
// Used to write the golden vectors under tests/golden/, to check the
// library's curvelet formation bit for bit, and as the timed CPU baseline of
// that stage.
//
// Drives form_curvelet_process as main_gpu_cpu.cpp:211-281 (and
// main_cpu_DP.cpp:208-281) do: the N x 4 edgel list transposed to column
// major, ctor(edgeinfo, H, W, rad, gap, dx, dt, token_len, max_k,
// max_size_to_group, centered, bidirectional), execute(), then
// get_output_size / get_output_arrary for each requested output type.  The
// reference allocates its output arrays with `new T[n]` (uninitialised) and
// writes only the cells it has values for; this harness zero-fills them, so
// unwritten cells read 0.
//
// usage: curvelet_ref <edges.f64> <N> <H> <W> <rad> <gap> <dx> <dt_rad>
//                     <token_len> <max_k> <style> <group> <types> <out.bin>
//   edges.f64: N x 4 doubles, row major [x, y, orientation, strength]
//   style: 0..3 as form_curvelet_main.cpp:42-44 maps cvlet_type to the
//          centered / bidirectional flags
//   types: digits of the output types to write, e.g. "01" or "012"
// out.bin: double t_execute (s), then per requested type:
//   int32 type, uint32 h, uint32 w, uint32 info_w,
//   h*w int32 chain (column major), h*info_w doubles info (column major)
#include <chrono>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "Array.hpp"
#include "form_curvelet_process.hpp"
#include "curvelet_utils.hpp"

int main(int argc, char **argv) {
    if (argc != 15) {
        std::fprintf(stderr, "usage: %s edges N H W rad gap dx dt token_len max_k style group types out\n", argv[0]);
        return 2;
    }
    const long n = std::atol(argv[2]);
    const unsigned H = (unsigned)std::atoi(argv[3]), W = (unsigned)std::atoi(argv[4]);
    const double rad = std::atof(argv[5]), gap = std::atof(argv[6]), dx = std::atof(argv[7]);
    const double dt = std::atof(argv[8]), token_len = std::atof(argv[9]), max_k = std::atof(argv[10]);
    const unsigned style = (unsigned)std::atoi(argv[11]), group = (unsigned)std::atoi(argv[12]);
    const char *types = argv[13];
    if (n < 0) return 2;

    std::vector<double> rows((size_t)n * 4);
    {
        std::FILE *f = std::fopen(argv[1], "rb");
        if (!f) return 2;
        const size_t got = std::fread(rows.data(), sizeof(double), rows.size(), f);
        std::fclose(f);
        if (got != rows.size()) return 2;
    }
    // main_gpu_cpu.cpp:231-236: the list goes in column major
    std::vector<double> cm((size_t)n * 4 + 1);
    for (long i = 0; i < n; ++i)
        for (int j = 0; j < 4; ++j) cm[(size_t)j * n + i] = rows[(size_t)i * 4 + j];
    arrayd edgeinfo;
    edgeinfo._data = cm.data();
    edgeinfo.set_h((unsigned)n);
    edgeinfo.set_w(4);

    const bool centered = style == 0 || style == 1;
    const bool bidirectional = style == 0 || style == 2;
    form_curvelet_process pro(edgeinfo, H, W, rad, gap, dx, dt, token_len, max_k, group, centered, bidirectional);
    const auto t0 = std::chrono::steady_clock::now();
    pro.execute();
    const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();

    std::FILE *out = std::fopen(argv[14], "wb");
    if (!out) return 2;
    std::fwrite(&secs, sizeof(double), 1, out);
    for (const char *t = types; *t; ++t) {
        const unsigned type = (unsigned)(*t - '0');
        if (type > 2) return 2;
        unsigned h = 0, w = 0;
        pro.get_output_size(h, w, type);
        const unsigned info_w = type == 0 ? 10u : type == 1 ? 1u : 12u;
        std::vector<int> chain((size_t)h * w + 1, 0);
        std::vector<double> info((size_t)h * info_w + 1, 0.0);
        arrayi ca;
        ca._data = chain.data();
        ca.set_h(h);
        ca.set_w(w);
        arrayd ia;
        ia._data = info.data();
        ia.set_h(h);
        ia.set_w(info_w);
        pro.get_output_arrary(ca, ia, type);
        ca._data = nullptr;     // the vectors own the storage
        ia._data = nullptr;
        const int32_t hdr[4] = {(int32_t)type, (int32_t)h, (int32_t)w, (int32_t)info_w};
        std::fwrite(hdr, sizeof(int32_t), 4, out);
        std::fwrite(chain.data(), sizeof(int), (size_t)h * w, out);
        std::fwrite(info.data(), sizeof(double), (size_t)h * info_w, out);
    }
    std::fclose(out);
    return 0;
}
