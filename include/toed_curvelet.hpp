// toed_curvelet.hpp -- header-only drop-in for the reference's curvelet
// formation class form_curvelet_process (curvelet/form_curvelet_process.hpp:
// 23-82) over the C ABI of libtoed_hip.so (toed_capi.h, toed_curvelet_*).
// Host code only; link with -ltoed_hip.
//
// The reference's drivers use it as main_gpu_cpu.cpp:238-281 does:
//
//   arrayd edgeinfo;  edgeinfo._data = cols;  edgeinfo.set_h(n);  edgeinfo.set_w(4);
//   form_curvelet_process pro(edgeinfo, H, W, nrad, gap, dx, dt, token_len, max_k,
//                             max_size_to_group, bCentered, bBidirectional);
//   pro.execute();
//   pro.get_output_size(out_h, out_w, output_type);
//   arrayi chain;  chain._data = out_chain;  chain.set_h(out_h);  chain.set_w(out_w);
//   arrayd info;   info._data = out_info;    info.set_h(out_h);   info.set_w(info_w);
//   pro.get_output_arrary(chain, info, output_type);
//
// The same code compiles against this header with `using namespace toed_cv;`
// (the arrays are column major, Array.hpp val(r, c) = _data[c * h + r], and
// the caller owns their storage).  Differences from the reference:
//   * results are the same bit for bit; cells the reference leaves unwritten
//     (chain columns past a chain's end, the type-1 info column) are written 0;
//   * execute() runs on every hardware thread (set_threads to change);
//   * output type 2 is written in full (the reference's build stops at the
//     assert in curvelet::set_output for that type);
//   * errors (bad arguments, output arrays of the wrong size) throw
//     std::runtime_error instead of printing to std::cerr.
#ifndef TOED_CURVELET_HPP
#define TOED_CURVELET_HPP

#include <stdexcept>
#include <string>

#include "toed_capi.h"

namespace toed_cv {

// A view of caller storage with the reference's Array<T> members that its
// drivers touch (Array.hpp: _data, h(), w(), set_h, set_w, val, set_val).
template <class T>
struct Array {
    T *_data = nullptr;
    unsigned h() const { return _h; }
    unsigned w() const { return _w; }
    void set_h(const unsigned &h) { _h = h; }
    void set_w(const unsigned &w) { _w = w; }
    T &val(const unsigned &r, const unsigned &c) const { return _data[(size_t)c * _h + r]; }
    void set_val(const unsigned &r, const unsigned &c, const T &v) { _data[(size_t)c * _h + r] = v; }

private:
    unsigned _h = 0, _w = 0;
};
typedef Array<double> arrayd;
typedef Array<int> arrayi;

class form_curvelet_process {
public:
    form_curvelet_process(arrayd &edgeinfo, const unsigned &nrows, const unsigned &ncols, const double &rad,
                          const double &gap, const double &dx, const double &dt, const double &token_len,
                          const double max_k, const unsigned max_size_to_goup, bool bCentered_grouping = false,
                          bool bBidirectional_grouping = false) {
        if (edgeinfo.w() < 4) throw std::runtime_error("form_curvelet_process: edgeinfo needs 4 columns");
        check(toed_curvelet_create(edgeinfo._data, edgeinfo.h(), nrows, ncols, rad, gap, dx, dt, token_len, max_k,
                                   max_size_to_goup, bCentered_grouping ? 1 : 0, bBidirectional_grouping ? 1 : 0,
                                   &h_),
              "toed_curvelet_create");
    }
    form_curvelet_process(const form_curvelet_process &) = delete;
    form_curvelet_process &operator=(const form_curvelet_process &) = delete;
    virtual ~form_curvelet_process() {
        if (h_) toed_curvelet_destroy(h_);
    }

    void set_threads(int nthreads) { nthreads_ = nthreads; }

    bool execute() {
        check(toed_curvelet_execute(h_, nthreads_), "toed_curvelet_execute");
        return true;
    }

    void get_output_size(unsigned &h, unsigned &w, unsigned output_type) {
        check(toed_curvelet_output_size(h_, output_type, &h, &w, nullptr), "toed_curvelet_output_size");
    }

    void get_output_arrary(arrayi &id_chain, arrayd &curvelet_info, unsigned output_type) {
        unsigned h = 0, w = 0, iw = 0;
        check(toed_curvelet_output_size(h_, output_type, &h, &w, &iw), "toed_curvelet_output_size");
        if (id_chain.h() != h || id_chain.w() != w || (output_type != 1 && (curvelet_info.h() != h ||
                                                                          curvelet_info.w() != iw)))
            throw std::runtime_error("form_curvelet_process: incorrect output array");
        check(toed_curvelet_output(h_, output_type, id_chain._data, curvelet_info._data), "toed_curvelet_output");
    }

private:
    toed_curvelets *h_ = nullptr;
    int nthreads_ = 0;

    static void check(int st, const char *what) {
        if (st != TOED_OK) throw std::runtime_error(std::string(what) + " failed (status " + std::to_string(st) + ")");
    }
};

}  // namespace toed_cv

#endif  // TOED_CURVELET_HPP
