// Command-line drop-in for the reference's TOED binary and its batch script.
//
//   toed_detect_cli <input_image> [nthreads] [gpu_id] [output_dir]
//       main_gpu_cpu.cpp:103-153: same arguments (nthreads is the reference's
//       OpenMP thread count for its CPU run and is accepted but unused),
//       output_dir default ./output_files (created), the edgel list written to
//       output_dir/data_final_output_gpu.txt as gpu_toed.hpp:452-478 writes it
//       (N rows of 4 values, each followed by a tab, precision 10).
//
//   toed_detect_cli --batch [--gpu N] [--out DIR] [--autotune] [--pgm-raw-offset] img...
//       run_gpu_batch.sh:152-193 in one process: every image in one batch
//       plan replayed as a hipGraph, each list written to
//       DIR/<stem>_edges_gpu.txt (the name the script gives it).
//
//   toed_detect_cli <input_image> ... --curvelets
//       also runs the curvelet formation the reference's driver runs on its
//       FP64 list when built with CurvelFormation (main_gpu_cpu.cpp:211-281):
//       the list widened to double, nrad 3.5, gap 1.5, dx 0.4, dt 15 degrees,
//       token_len 1, max_k 0.3, style 2, groups of 4, output type 0, written
//       to output_dir/chain.txt and info.txt as main_gpu_cpu.cpp:69-97 writes
//       them (default stream format, a tab after every value).
//
// sigma 2, threshold 2, border 10 (main_gpu_cpu.cpp:158-159, cpu_toed.cpp:
// 400-403).  Images: 8-bit PGM (P5) or PNG via toed_image_load; with
// --pgm-raw-offset a PGM is read as the reference's no-OpenCV path reads it
// (every pixel one byte early).  Exit status 0 on success.
#include <hip/hip_runtime.h>

#include <sys/stat.h>

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

#include "toed_capi.h"
#include "toed_curvelet.hpp"

namespace {

bool g_curvelets = false;

const float kSigma = 2.0f, kThresh = 2.0f;
const int kBorder = 10;

bool mkdir_p(const std::string &dir) {
    if (dir.empty()) return false;
    std::string cur;
    for (size_t i = 0; i <= dir.size(); ++i) {
        if (i == dir.size() || dir[i] == '/') {
            if (!cur.empty() && cur != "/") {
                struct stat st;
                if (stat(cur.c_str(), &st) != 0) {
                    if (mkdir(cur.c_str(), 0755) != 0) return false;
                } else if (!S_ISDIR(st.st_mode)) {
                    return false;
                }
            }
        }
        if (i < dir.size()) cur.push_back(dir[i]);
    }
    return true;
}

std::string join(const std::string &dir, const std::string &name) {
    std::string p = dir;
    if (!p.empty() && p.back() != '/') p.push_back('/');
    return p + name;
}

std::string stem_of(const std::string &path) {
    const size_t s = path.find_last_of('/');
    std::string base = s == std::string::npos ? path : path.substr(s + 1);
    const size_t d = base.find_last_of('.');
    return d == std::string::npos || d == 0 ? base : base.substr(0, d);
}

// gpu_toed.hpp:452-478: precision 10, a tab after every value, a newline per row
bool write_list(const std::string &path, const std::vector<float> &rows, long long n) {
    std::ofstream out(path);
    if (!out.is_open()) return false;
    out.precision(10);
    for (long long i = 0; i < n; ++i) {
        for (int j = 0; j < 4; ++j) out << rows[(size_t)i * 4 + j] << "\t";
        out << "\n";
    }
    return (bool)out;
}

// main_gpu_cpu.cpp:69-97 _write_array_to_file_colmajor: column-major data,
// one row per line, a tab after every value, default stream format
template <class T>
bool write_colmajor(const std::string &path, const T *data, unsigned h, unsigned w) {
    std::ofstream out(path);
    if (!out.is_open()) return false;
    for (unsigned i = 0; i < h; ++i) {
        for (unsigned j = 0; j < w; ++j) out << data[i + (size_t)h * j] << "\t";
        out << "\n";
    }
    return (bool)out;
}

// main_gpu_cpu.cpp:211-281: curvelets of the list, chain.txt and info.txt
bool write_curvelets(const std::string &out_dir, const std::vector<float> &rows, long long n, int H, int W) {
    std::vector<double> cols((size_t)n * 4 + 1);
    for (long long i = 0; i < n; ++i)
        for (int j = 0; j < 4; ++j) cols[(size_t)j * n + i] = (double)rows[(size_t)i * 4 + j];
    toed_cv::arrayd edgeinfo;
    edgeinfo._data = cols.data();
    edgeinfo.set_h((unsigned)n);
    edgeinfo.set_w(4);
    const double pi = 3.14159265358979323846;
    const unsigned style = 2, group = 4, output_type = 0;
    toed_cv::form_curvelet_process pro(edgeinfo, (unsigned)H, (unsigned)W, 3.5, 1.5, 0.4, (15.0 / 180.0) * pi, 1, 0.3,
                                       group, style == 0 || style == 1, style == 0 || style == 2);
    const auto t0 = std::chrono::steady_clock::now();
    pro.execute();
    const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    unsigned out_h = 0, out_w = 0;
    pro.get_output_size(out_h, out_w, output_type);
    std::vector<int> chain((size_t)out_h * out_w + 1);
    std::vector<double> info((size_t)out_h * 10 + 1);
    toed_cv::arrayi ca;
    ca._data = chain.data();
    ca.set_h(out_h);
    ca.set_w(out_w);
    toed_cv::arrayd ia;
    ia._data = info.data();
    ia.set_h(out_h);
    ia.set_w(10);
    pro.get_output_arrary(ca, ia, output_type);
    std::cout << "Curvelet formation: " << out_h << " curvelets in " << ms << " ms" << std::endl;
    return write_colmajor(join(out_dir, "chain.txt"), chain.data(), out_h, out_w) &&
           write_colmajor(join(out_dir, "info.txt"), info.data(), out_h, 10u);
}

struct Image {
    std::string path;
    int H = 0, W = 0;
    unsigned char *d_img = nullptr;
    float *d_edges = nullptr;
    long long capacity = 0;
};

int load_to_device(Image &im, int raw_offset) {
    unsigned char *px = nullptr;
    int st = toed_image_load(im.path.c_str(), raw_offset, &px, &im.H, &im.W);
    if (st) return st;
    const size_t n = (size_t)im.H * im.W;
    // the reference's list buffer holds every interp pixel: (2H)(2W) rows
    im.capacity = 4LL * im.H * im.W;
    if (hipMalloc((void **)&im.d_img, n) != hipSuccess ||
        hipMalloc((void **)&im.d_edges, sizeof(float) * 4 * (size_t)im.capacity) != hipSuccess) {
        toed_image_free(px);
        return TOED_ERR_ALLOC;
    }
    const hipError_t e = hipMemcpy(im.d_img, px, n, hipMemcpyHostToDevice);
    toed_image_free(px);
    return e == hipSuccess ? TOED_OK : TOED_ERR_HIP;
}

toed_image_t describe(const Image &im) {
    toed_image_t t;
    std::memset(&t, 0, sizeof t);
    t.in_type = TOED_U8;
    t.d_img = im.d_img;
    t.H = im.H;
    t.W = im.W;
    t.pitch = im.W;
    t.sigma = kSigma;
    t.thresh = kThresh;
    t.border = kBorder;
    t.d_edges = im.d_edges;
    t.capacity = im.capacity;
    return t;
}

void release(Image &im) {
    if (im.d_img) (void)hipFree(im.d_img);
    if (im.d_edges) (void)hipFree(im.d_edges);
    im.d_img = nullptr;
    im.d_edges = nullptr;
}

int run_single(int argc, char **argv) {
    Image im;
    im.path = argv[1];
    {
        std::ifstream probe(im.path, std::ios::binary);
        if (!probe.is_open()) {   // main_gpu_cpu.cpp:114-120 (exit status 0 there too)
            std::cout << "File " << im.path << " not found in directory." << std::endl;
            return 0;
        }
    }
    const int gpu_id = argc > 3 ? std::atoi(argv[3]) : 0;
    const std::string out_dir = argc > 4 ? argv[4] : "./output_files";
    if (!mkdir_p(out_dir)) {
        std::cerr << "Error: failed to create output directory: " << out_dir << std::endl;
        return 1;
    }
    std::cout << "Output directory: " << out_dir << std::endl;
    const char *raw = std::getenv("TOED_PGM_RAW_OFFSET");
    if (hipSetDevice(gpu_id) != hipSuccess) { std::cerr << "Error: no GPU " << gpu_id << std::endl; return 1; }
    int st = load_to_device(im, raw && raw[0] == '1');
    if (st) { std::cerr << "Error: cannot read image " << im.path << " (status " << st << ")" << std::endl; return 1; }
    toed_ctx *ctx = nullptr;
    if ((st = toed_create(gpu_id, &ctx))) { std::cerr << "Error: toed_create " << st << std::endl; return 1; }
    const toed_image_t t = describe(im);
    long long n = 0;
    st = toed_detect_sync(ctx, &t, &n);
    if (st) { std::cerr << "Error: toed_detect_sync " << st << std::endl; toed_destroy(ctx); release(im); return 1; }
    std::vector<float> rows((size_t)n * 4);
    if (n && hipMemcpy(rows.data(), im.d_edges, rows.size() * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess) {
        toed_destroy(ctx); release(im); return 1;
    }
    float ms_tiles = 0.f, ms_final = 0.f;
    (void)toed_last_times(ctx, &ms_tiles, &ms_final);
    std::cout << "Image " << im.H << " x " << im.W << ": GPU detection " << (ms_tiles + ms_final) << " ms" << std::endl;
    std::cout << "Number of GPU edges = " << n << std::endl;
    const std::string path = join(out_dir, "data_final_output_gpu.txt");
    std::cout << "writing data to a file " << path << " ..." << std::endl;
    bool ok = write_list(path, rows, n);
    toed_destroy(ctx);
    release(im);
    if (ok && g_curvelets) {
        try {
            ok = write_curvelets(out_dir, rows, n, im.H, im.W);
        } catch (const std::exception &e) {
            std::cerr << "Error: " << e.what() << std::endl;
            ok = false;
        }
    }
    return ok ? 0 : 1;
}

int run_batch(int argc, char **argv) {
    int gpu_id = 0, raw_offset = 0;
    bool tune = false;
    std::string out_dir = "./output_files";
    std::vector<Image> ims;
    for (int a = 2; a < argc; ++a) {
        const std::string s = argv[a];
        if (s == "--gpu" && a + 1 < argc) gpu_id = std::atoi(argv[++a]);
        else if (s == "--out" && a + 1 < argc) out_dir = argv[++a];
        else if (s == "--autotune") tune = true;
        else if (s == "--pgm-raw-offset") raw_offset = 1;
        else { Image im; im.path = s; ims.push_back(im); }
    }
    if (ims.empty()) { std::cerr << "Error: no images" << std::endl; return 1; }
    if (!mkdir_p(out_dir)) { std::cerr << "Error: failed to create output directory: " << out_dir << std::endl; return 1; }
    if (hipSetDevice(gpu_id) != hipSuccess) { std::cerr << "Error: no GPU " << gpu_id << std::endl; return 1; }
    int failed = 0;
    std::vector<Image> ok;
    for (Image &im : ims) {
        const int st = load_to_device(im, raw_offset);
        if (st) {
            std::cerr << "  FAILED: cannot read " << im.path << " (status " << st << ")" << std::endl;
            release(im);
            ++failed;
        } else {
            ok.push_back(im);
        }
    }
    int rc = failed ? 1 : 0;
    toed_ctx *ctx = nullptr;
    if (ok.empty() || toed_create(gpu_id, &ctx)) { for (Image &im : ok) release(im); return 1; }
    const int n = (int)ok.size();
    std::vector<toed_image_t> t(n);
    for (int m = 0; m < n; ++m) t[m] = describe(ok[m]);
    toed_tune_t best{16, 0, 0, 0.f};
    if (tune && toed_batch_autotune(ctx, n, t.data(), 3, &best, nullptr, 0, nullptr) == TOED_OK)
        std::cout << "autotuned: " << best.tile_rows << " tile rows, "
                  << (best.persistent ? "persistent grid" : "one workgroup per tile") << std::endl;
    int *d_cf = nullptr;
    toed_batch *b = nullptr;
    int st = hipMalloc((void **)&d_cf, sizeof(int) * 2 * n) == hipSuccess ? TOED_OK : TOED_ERR_ALLOC;
    if (!st) st = toed_batch_plan(ctx, n, t.data(), &b);
    if (!st) st = toed_batch_set_persistent(b, best.persistent);
    if (!st) st = toed_batch_set_grid(b, best.wg_per_cu);
    const auto t0 = std::chrono::steady_clock::now();
    if (!st) st = toed_batch_graph_run(b, d_cf, d_cf + n);
    std::vector<int> cf(2 * n, 0);
    if (!st && hipMemcpy(cf.data(), d_cf, sizeof(int) * 2 * n, hipMemcpyDeviceToHost) != hipSuccess) st = TOED_ERR_HIP;
    const double sec = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (st) {
        std::cerr << "Error: batch detection failed (status " << st << ")" << std::endl;
        rc = 1;
    } else {
        std::cout << n << " images in one plan: " << sec * 1e3 << " ms (first run, including graph capture)"
                  << std::endl;
        for (int m = 0; m < n; ++m) {
            const long long cnt = cf[m];
            const std::string path = join(out_dir, stem_of(ok[m].path) + "_edges_gpu.txt");
            if (cf[n + m] || cnt > ok[m].capacity) {
                std::cerr << "  FAILED: " << ok[m].path << " (flags " << cf[n + m] << ")" << std::endl;
                rc = 1;
                continue;
            }
            std::vector<float> rows((size_t)cnt * 4);
            if (cnt && hipMemcpy(rows.data(), ok[m].d_edges, rows.size() * sizeof(float), hipMemcpyDeviceToHost) !=
                           hipSuccess) {
                rc = 1;
                continue;
            }
            if (!write_list(path, rows, cnt)) { rc = 1; continue; }
            std::cout << "  " << ok[m].path << ": " << cnt << " edges -> " << path << std::endl;
        }
    }
    if (b) toed_batch_destroy(b);
    if (d_cf) (void)hipFree(d_cf);
    toed_destroy(ctx);
    for (Image &im : ok) release(im);
    return rc;
}

}  // namespace

int main(int argc, char **argv) {
    // --curvelets may appear anywhere in the single-image form
    int m = 1;
    for (int a = 1; a < argc; ++a) {
        if (!std::strcmp(argv[a], "--curvelets"))
            g_curvelets = true;
        else
            argv[m++] = argv[a];
    }
    argc = m;
    if (argc < 2) {
        std::cout << "Usage: " << argv[0] << " <input_image> [nthreads] [gpu_id] [output_dir] [--curvelets]\n"
                  << "       " << argv[0]
                  << " --batch [--gpu N] [--out DIR] [--autotune] [--pgm-raw-offset] image...  "
                     "(writes DIR/<stem>_edges_gpu.txt)" << std::endl;
        return 0;
    }
    if (!std::strcmp(argv[1], "--batch")) return run_batch(argc, argv);
    return run_single(argc, argv);
}
