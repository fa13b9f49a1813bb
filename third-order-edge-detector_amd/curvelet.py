"""Curvelet formation on an edgel list (host code in libtoed_hip.so).

The consumer the reference runs on its FP64 detector output
(main_gpu_cpu.cpp:211-281, main_cpu_DP.cpp:208-281):
``curvelet/form_curvelet_process.hpp`` groups edgels into curvelets, links
them, extracts curve fragments and fits poly-arcs.  ``CurveletFormation``
mirrors that class (constructor, ``execute``, ``get_output_size``,
``get_output_arrary``); ``form_curvelets`` mirrors
``curvelet/form_curvelet_main.cpp:32-105`` ``curvelet_formation`` (row-major
N x 4 list in, chain and info arrays out) with the reference drivers'
settings as defaults: nrad 3.5, gap 1.5, dx 0.4, dt 15 degrees, token_len 1,
max_k 0.3, style 2 (anchor-leading bidirectional), groups of 4, output
type 0 (main_gpu_cpu.cpp:217-228).

Outputs are returned as H x W numpy arrays with the reference's values
(``chain``: 1-based edgel ids, the curvelet map's first column the 1-based
anchor id; ``info``: 10 / 1 / 12 columns by output type).  Cells the
reference leaves unwritten are 0.  Results are bit-identical to the
reference's compiled code and do not depend on the thread count.
"""
import ctypes
import math

import numpy as np

from ._lib import check, lib

OUTPUT_CURVELETS, OUTPUT_FRAGMENTS, OUTPUT_POLYARCS = 0, 1, 2
INFO_WIDTH = {0: 10, 1: 1, 2: 12}


def style_flags(curvelet_style):
    """form_curvelet_main.cpp:92: style 0 anchor-centred bidirectional,
    1 anchor-centred, 2 anchor-leading bidirectional, 3 ENO style."""
    s = int(curvelet_style)
    if s not in (0, 1, 2, 3):
        raise ValueError("curvelet_style must be 0..3")
    return s in (0, 1), s in (0, 2)


def _edges_numpy(edges):
    if hasattr(edges, "detach"):            # a torch tensor (CPU or device)
        edges = edges.detach().to("cpu").double().numpy()
    e = np.asarray(edges, dtype=np.float64)
    if e.ndim != 2 or (e.shape[0] and e.shape[1] < 4):
        raise ValueError("edges must be an N x 4 (or wider) array")
    return e


class CurveletFormation:
    """form_curvelet_process (curvelet/form_curvelet_process.hpp:23-82)."""

    def __init__(self, edges, nrows, ncols, rad=3.5, gap=1.5, dx=0.4, dt=15.0 / 180.0 * math.pi, token_len=1.0,
                 max_k=0.3, max_size_to_group=4, centered=False, bidirectional=True):
        e = _edges_numpy(edges)
        n = int(e.shape[0])
        # column major, as main_gpu_cpu.cpp:231-236 hands the list over
        cols = np.ascontiguousarray(e[:, :4].T) if n else np.zeros((4, 1))
        h = ctypes.c_void_p()
        check(lib().toed_curvelet_create(cols.ctypes.data, n, int(nrows), int(ncols), float(rad), float(gap),
                                         float(dx), float(dt), float(token_len), float(max_k),
                                         int(max_size_to_group), 1 if centered else 0, 1 if bidirectional else 0,
                                         ctypes.byref(h)), "toed_curvelet_create")
        self._h = h
        self.n_edgels = n

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            lib().toed_curvelet_destroy(h)
            self._h = None

    def execute(self, nthreads=0):
        check(lib().toed_curvelet_execute(self._h, int(nthreads)), "toed_curvelet_execute")
        return True

    def get_output_size(self, output_type=0):
        """(h, w) of the chain array for the output type."""
        oh, ow, iw = ctypes.c_uint(), ctypes.c_uint(), ctypes.c_uint()
        check(lib().toed_curvelet_output_size(self._h, int(output_type), ctypes.byref(oh), ctypes.byref(ow),
                                              ctypes.byref(iw)), "toed_curvelet_output_size")
        return oh.value, ow.value

    def get_output_arrary(self, output_type=0):
        """(chain, info): int32 h x w and float64 h x info_w arrays."""
        h, w = self.get_output_size(output_type)
        iw = INFO_WIDTH[int(output_type)]
        chain = np.empty((max(w, 1), max(h, 1)), dtype=np.int32)     # column major storage
        info = np.empty((iw, max(h, 1)), dtype=np.float64)
        check(lib().toed_curvelet_output(self._h, int(output_type), chain.ctypes.data, info.ctypes.data),
              "toed_curvelet_output")
        return chain[:w, :h].T, info[:, :h].T

    get_output_array = get_output_arrary


def form_curvelets(edges, height, width, nrad=3.5, gap=1.5, dx=0.4, dt=15.0 / 180.0 * math.pi, token_len=1.0,
                   max_k=0.3, curvelet_style=2, max_size_to_group=4, output_type=0, nthreads=0):
    """curvelet_formation (form_curvelet_main.cpp:32-105): row-major N x 4
    edgel list in, (chain, info) out."""
    centered, bidirectional = style_flags(curvelet_style)
    cf = CurveletFormation(edges, height, width, nrad, gap, dx, dt, token_len, max_k, max_size_to_group, centered,
                           bidirectional)
    cf.execute(nthreads)
    return cf.get_output_arrary(output_type)
