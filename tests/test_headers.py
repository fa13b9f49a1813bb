"""The C++ drop-in headers (include/toed_gpu.hpp, include/toed_subpix_gpu.hpp,
include/toed_curvelet.hpp) compile as a host translation unit that instantiates the classes the way the
reference's drivers do (CPU only: nothing is run)."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

SRC = r'''
#include "toed_gpu.hpp"
#include "toed_subpix_gpu.hpp"
#include "toed_curvelet.hpp"
int use_classes(int H, int W, const double *E, double *out) {
    // main_subpix_to_correction.cpp:279-289, class name changed
    SubpixelTOCorrectionGPU<double> to_corr(H, W, 0.1, 2.0, 8, 0);
    to_corr.set_output_dir("./out");
    to_corr.set_neighbor_radius(2.0);
    to_corr.set_edgemap(E);
    to_corr.estimate_orientation();
    int n = to_corr.run(out);
    to_corr.write_array_to_file("TO_corrected_edges.txt", out, n, 4);
    to_corr.write_array_to_file("TO_edgemap.txt", to_corr.TO_edgemap, H, W);
    // main_gpu_cpu.cpp:158-190 with the GPU class
    ThirdOrderEdgeDetectionGPU<float> det(0, H, W, 17, 2);
    // main_gpu_cpu.cpp:200-209: the double instantiation
    ThirdOrderEdgeDetectionGPU<double> det64(0, H, W, 17, 2);
    det64.convolve_img();
    return n + det.edge_pt_list_idx + det64.non_maximum_suppresion(out);
}
using namespace toed_cv;
int use_curvelets(int height, int width, double *TOED_edges_cm, int edge_num, int *out_chain, double *out_info) {
    // main_gpu_cpu.cpp:238-276, unchanged apart from the namespace
    double nrad = 3.5, gap = 1.5, dx = 0.4, dt = (15.0 / 180.0) * 3.14159265358979323846, token_len = 1,
           max_k = 0.3;
    unsigned curvelet_style = 2, max_size_to_group = 4, output_type = 0;
    arrayd edgeinfo;
    edgeinfo._data = TOED_edges_cm;
    edgeinfo.set_h(edge_num);
    edgeinfo.set_w(4);
    unsigned cvlet_type = curvelet_style;
    bool bCentered_grouping = cvlet_type == 0 || cvlet_type == 1;
    bool bBidirectional_grouping = cvlet_type == 0 || cvlet_type == 2;
    form_curvelet_process curvelet_pro(edgeinfo, unsigned(height), unsigned(width), nrad, gap, dx, dt, token_len,
                                       max_k, max_size_to_group, bCentered_grouping, bBidirectional_grouping);
    curvelet_pro.execute();
    unsigned out_h, out_w, info_w = 10;
    curvelet_pro.get_output_size(out_h, out_w, output_type);
    arrayi chain;
    chain._data = out_chain;
    chain.set_h(out_h);
    chain.set_w(out_w);
    arrayd info;
    info._data = out_info;
    info.set_h(out_h);
    info.set_w(info_w);
    curvelet_pro.get_output_arrary(chain, info, output_type);
    return (int)(chain.w() + info.h());
}
'''


def test_headers_compile(tmp_path):
    gxx = shutil.which("g++")
    if not gxx or not os.path.isdir("/opt/rocm/include"):
        pytest.skip("no host compiler / ROCm headers")
    src = tmp_path / "use.cpp"
    src.write_text(SRC)
    cmd = [gxx, "-std=c++17", "-fsyntax-only", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include",
           "-I" + os.path.join(ROOT, "include"), str(src)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    assert res.returncode == 0, res.stderr[-3000:]
