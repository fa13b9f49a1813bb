// Image files for the command-line drop-in: 8-bit grey PGM (P5) and PNG.
//
// The reference reads its input in main_gpu_cpu.cpp:114-132: with OpenCV by
// cv::imread(IMREAD_GRAYSCALE), without it by `infile >> type >> width >>
// height >> intensity` followed by one infile.get() per pixel
// (gpu_toed.hpp:189-198, cpu_toed.cpp:99-106).  The latter leaves the single
// whitespace byte that ends a PGM header in the stream, so every pixel it
// stores is the file's previous byte (the first one is that whitespace).
// toed_image_load decodes the file as the format defines it; with
// pgm_raw_offset = 1 it reproduces the no-OpenCV read instead, byte for byte.
//
// Host code only (zlib for PNG's deflate streams).
#include <zlib.h>

#include <cctype>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "toed_capi.h"

namespace {

bool read_file(const char *path, std::vector<unsigned char> &buf) {
    FILE *f = std::fopen(path, "rb");
    if (!f) return false;
    unsigned char tmp[1 << 16];
    size_t k;
    while ((k = std::fread(tmp, 1, sizeof tmp, f)) > 0) buf.insert(buf.end(), tmp, tmp + k);
    std::fclose(f);
    return true;
}

// ---- PGM ------------------------------------------------------------------
// P5 header tokens are separated by whitespace; '#' starts a comment that runs
// to the end of the line; exactly one whitespace byte follows the maxval.
bool pgm_token(const std::vector<unsigned char> &b, size_t &p, bool comments, long &v) {
    for (;;) {
        while (p < b.size() && std::isspace(b[p])) ++p;
        if (comments && p < b.size() && b[p] == '#') {
            while (p < b.size() && b[p] != '\n' && b[p] != '\r') ++p;
            continue;
        }
        break;
    }
    if (p >= b.size() || !std::isdigit(b[p])) return false;
    v = 0;
    while (p < b.size() && std::isdigit(b[p])) {
        v = v * 10 + (b[p] - '0');
        if (v > (1L << 30)) return false;
        ++p;
    }
    return true;
}

int load_pgm(const std::vector<unsigned char> &b, int raw_offset, std::vector<unsigned char> &px, int &H, int &W) {
    size_t p = 2;
    long w, h, mx;
    const bool comments = !raw_offset;   // operator>> knows no comments
    if (!pgm_token(b, p, comments, w) || !pgm_token(b, p, comments, h) || !pgm_token(b, p, comments, mx))
        return TOED_ERR_ARG;
    if (w <= 0 || h <= 0 || mx <= 0 || mx > 255) return TOED_ERR_ARG;
    // format: one whitespace byte, then the samples; raw_offset: the samples
    // start at that whitespace byte (the reference's operator>> stops before it)
    if (!raw_offset) {
        if (p >= b.size() || !std::isspace(b[p])) return TOED_ERR_ARG;
        ++p;
    }
    const size_t n = (size_t)w * (size_t)h;
    if (b.size() - p < n) return TOED_ERR_ARG;
    px.assign(b.begin() + p, b.begin() + p + n);
    H = (int)h;
    W = (int)w;
    return TOED_OK;
}

// ---- PNG ------------------------------------------------------------------
uint32_t be32(const unsigned char *q) {
    return ((uint32_t)q[0] << 24) | ((uint32_t)q[1] << 16) | ((uint32_t)q[2] << 8) | q[3];
}

int paeth(int a, int b, int c) {
    const int p = a + b - c, pa = std::abs(p - a), pb = std::abs(p - b), pc = std::abs(p - c);
    if (pa <= pb && pa <= pc) return a;
    return pb <= pc ? b : c;
}

// Non-interlaced 8- or 16-bit grey, grey + alpha, RGB or RGBA.  Alpha is
// dropped, 16-bit samples keep their high byte, colour goes to grey with the
// ITU-R BT.601 weights (0.299, 0.587, 0.114) in 14-bit fixed point, rounded.
int load_png(const std::vector<unsigned char> &b, std::vector<unsigned char> &px, int &H, int &W) {
    static const unsigned char sig[8] = {137, 80, 78, 71, 13, 10, 26, 10};
    if (b.size() < 8 || std::memcmp(b.data(), sig, 8) != 0) return TOED_ERR_ARG;
    size_t p = 8;
    uint32_t w = 0, h = 0;
    int depth = 0, ctype = -1, interlace = 0;
    std::vector<unsigned char> z;
    bool ihdr = false, iend = false;
    while (!iend && p + 12 <= b.size()) {
        const uint32_t len = be32(&b[p]);
        if (len > b.size() - p - 12) return TOED_ERR_ARG;
        const unsigned char *type = &b[p + 4], *d = &b[p + 8];
        if (!std::memcmp(type, "IHDR", 4)) {
            if (len != 13) return TOED_ERR_ARG;
            w = be32(d);
            h = be32(d + 4);
            depth = d[8];
            ctype = d[9];
            if (d[10] != 0 || d[11] != 0) return TOED_ERR_ARG;   // deflate, adaptive filtering
            interlace = d[12];
            ihdr = true;
        } else if (!std::memcmp(type, "IDAT", 4)) {
            z.insert(z.end(), d, d + len);
        } else if (!std::memcmp(type, "IEND", 4)) {
            iend = true;
        }
        p += 12 + (size_t)len;
    }
    if (!ihdr || w == 0 || h == 0 || w > (1u << 20) || h > (1u << 20) || interlace != 0) return TOED_ERR_ARG;
    int ch;
    switch (ctype) {
        case 0: ch = 1; break;
        case 2: ch = 3; break;
        case 4: ch = 2; break;
        case 6: ch = 4; break;
        default: return TOED_ERR_ARG;                            // palette or invalid
    }
    if (depth != 8 && depth != 16) return TOED_ERR_ARG;
    const int bpp = ch * depth / 8;
    const size_t stride = (size_t)w * bpp;
    std::vector<unsigned char> raw((stride + 1) * h);
    uLongf out_len = (uLongf)raw.size();
    if (uncompress(raw.data(), &out_len, z.data(), (uLong)z.size()) != Z_OK || out_len != raw.size())
        return TOED_ERR_ARG;
    // undo the per-row filters in place (each row's filter byte, then its bytes)
    std::vector<unsigned char> prev(stride, 0), cur(stride);
    px.resize((size_t)w * h);
    for (uint32_t y = 0; y < h; ++y) {
        const unsigned char f = raw[y * (stride + 1)];
        const unsigned char *s = &raw[y * (stride + 1) + 1];
        for (size_t x = 0; x < stride; ++x) {
            const int a = x >= (size_t)bpp ? cur[x - bpp] : 0, up = prev[x];
            const int c = x >= (size_t)bpp ? prev[x - bpp] : 0;
            int v = s[x];
            switch (f) {
                case 0: break;
                case 1: v += a; break;
                case 2: v += up; break;
                case 3: v += (a + up) >> 1; break;
                case 4: v += paeth(a, up, c); break;
                default: return TOED_ERR_ARG;
            }
            cur[x] = (unsigned char)v;
        }
        const int step = depth / 8;   // a sample's high byte comes first
        for (uint32_t x = 0; x < w; ++x) {
            const unsigned char *q = &cur[(size_t)x * bpp];
            int g;
            if (ch <= 2) {
                g = q[0];
            } else {
                const int r = q[0], gr = q[step], bl = q[2 * step];
                g = (r * 4899 + gr * 9617 + bl * 1868 + (1 << 13)) >> 14;
            }
            px[(size_t)y * w + x] = (unsigned char)g;
        }
        prev.swap(cur);
    }
    H = (int)h;
    W = (int)w;
    return TOED_OK;
}

}  // namespace

extern "C" {

int toed_image_load(const char *path, int pgm_raw_offset, unsigned char **pixels, int *H, int *W) {
    if (!path || !pixels || !H || !W) return TOED_ERR_ARG;
    *pixels = nullptr;
    std::vector<unsigned char> b, px;
    if (!read_file(path, b)) return TOED_ERR_ARG;
    int h = 0, w = 0, st;
    if (b.size() >= 2 && b[0] == 'P' && b[1] == '5')
        st = load_pgm(b, pgm_raw_offset, px, h, w);
    else
        st = load_png(b, px, h, w);
    if (st) return st;
    unsigned char *out = (unsigned char *)std::malloc(px.size());
    if (!out) return TOED_ERR_ALLOC;
    std::memcpy(out, px.data(), px.size());
    *pixels = out;
    *H = h;
    *W = w;
    return TOED_OK;
}

void toed_image_free(unsigned char *pixels) { std::free(pixels); }

}  // extern "C"
