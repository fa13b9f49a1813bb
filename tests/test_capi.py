"""The C-ABI library loads and exports every function include/*.h declares
(CPU only: no compute call is made)."""
import ctypes
import glob
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = {}
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        src = open(h).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        for m in re.finditer(r"^\s*(?:int|void\s*\*|void|long long|double|float)\s+(toed_\w+)\s*\(", src, re.M):
            names[m.group(1)] = os.path.basename(h)
    return names


def test_header_declares_entry_points():
    names = declared_functions()
    for must in ["toed_create", "toed_destroy", "toed_detect", "toed_detect_sync", "toed_batch_plan",
                 "toed_batch_run", "toed_batch_graph_run", "toed_gpu_convolve_f32", "toed_gpu_nms_f32",
                 "toed_gpu_convolve_f64", "toed_gpu_nms_f64"]:
        assert must in names


def test_library_exports_every_declared_symbol(pkg):
    from importlib import import_module
    L = import_module(pkg.__name__ + "._lib")
    lib = ctypes.CDLL(L.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing
    # the Python binding covers the same set
    bound = {s[0] for s in L.SIGNATURES}
    assert set(declared_functions()) <= bound, set(declared_functions()) - bound


def test_sigma_radius_host_only(pkg):
    assert pkg.sigma_radius(2.0) == (8, 9)
    assert pkg.sigma_radius(1.5) == (6, 7)
    assert pkg.sigma_radius(1.0) == (4, 5)
    for bad in [0.0, -1.0, 2.5, float("nan")]:
        with pytest.raises(pkg.ToedError):
            pkg.sigma_radius(bad)


def test_subpix_radius_host_only(pkg):
    assert pkg.subpix_radius(2.0) == 8
    assert pkg.subpix_radius(1.5) == 6
    assert pkg.subpix_radius(3.0) == 12
    for bad in [0.0, -1.0, 3.1, float("nan")]:
        with pytest.raises(pkg.ToedError):
            pkg.subpix_radius(bad)


def test_autograd_wrapper_host_logic():
    """The differentiable wrapper's host helpers (no GPU): sigma's host value
    from a float or a host tensor (no autograd warning), and the plan cache
    keyed by shape and parameters with at most two plans."""
    import importlib
    import warnings
    torch = pytest.importorskip("torch")
    ag = importlib.import_module("third-order-edge-detector_amd.autograd")
    t = torch.tensor(1.5, dtype=torch.float64, requires_grad=True)
    with warnings.catch_warnings():
        warnings.simplefilter("error")
        assert ag._sigma_value(t) == 1.5
    assert ag._sigma_value(2.0) == 2.0
    assert callable(ag._plan_for) and ag._MAX_PLANS == 2
