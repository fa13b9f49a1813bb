"""ctypes binding of libtoed_hip.so (the C ABI declared in include/toed_capi.h).

The library is the only compute path: if it is missing, every entry point
raises instead of falling back to anything else.
"""
import ctypes
import os

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libtoed_hip.so")

TOED_U8, TOED_F32 = 0, 1
FLAG_CAPACITY, FLAG_STAGING = 1, 2
TOED_SUBPIX_MAX_INTERP = 4     # toed_capi.h

STATUS = {
    0: "TOED_OK", 1: "TOED_ERR_ARG", 2: "TOED_ERR_SIGMA", 3: "TOED_ERR_ALLOC", 4: "TOED_ERR_HIP",
    5: "TOED_ERR_CAPACITY", 6: "TOED_ERR_STAGING", 7: "TOED_ERR_NOT_BUILT",
}


class ToedError(RuntimeError):
    def __init__(self, code, what=""):
        self.code = code
        super().__init__("%s failed: %s (%d)" % (what or "toed call", STATUS.get(code, "?"), code))


class toed_image_t(ctypes.Structure):
    _fields_ = [
        ("in_type", ctypes.c_int),
        ("d_img", ctypes.c_void_p),
        ("H", ctypes.c_int), ("W", ctypes.c_int), ("pitch", ctypes.c_int),
        ("sigma", ctypes.c_float), ("thresh", ctypes.c_float), ("border", ctypes.c_int),
        ("d_edges", ctypes.c_void_p),
        ("capacity", ctypes.c_longlong),
        ("d_index", ctypes.c_void_p),
    ]


class toed_grad_t(ctypes.Structure):
    _fields_ = [
        ("in_type", ctypes.c_int),
        ("d_img", ctypes.c_void_p),
        ("H", ctypes.c_int), ("W", ctypes.c_int), ("pitch", ctypes.c_int),
        ("sigma", ctypes.c_float),
        ("d_index", ctypes.c_void_p),
        ("d_count", ctypes.c_void_p),
        ("max_edgels", ctypes.c_int),
        ("d_gout", ctypes.c_void_p),
        ("d_gimg", ctypes.c_void_p),
        ("gpitch", ctypes.c_int),
        ("d_gsigma", ctypes.c_void_p),
    ]


class toed_grad_out_t(ctypes.Structure):
    _fields_ = [
        ("d_gout", ctypes.c_void_p),
        ("d_gimg", ctypes.c_void_p),
        ("gpitch", ctypes.c_int),
        ("d_gsigma", ctypes.c_void_p),
    ]


class toed_subpix_map_t(ctypes.Structure):
    _fields_ = [
        ("d_map", ctypes.c_void_p),
        ("d_orient", ctypes.c_void_p),
        ("H", ctypes.c_int), ("W", ctypes.c_int), ("pitch", ctypes.c_int), ("opitch", ctypes.c_int),
        ("thresh", ctypes.c_double),
        ("d_edges", ctypes.c_void_p),
        ("capacity", ctypes.c_longlong),
        ("d_index", ctypes.c_void_p),
        ("d_to_edgemap", ctypes.c_void_p),
        ("d_to_orientation", ctypes.c_void_p),
        ("tpitch", ctypes.c_int),
    ]


class toed_subpix_grad_t(ctypes.Structure):
    _fields_ = [
        ("d_gout", ctypes.c_void_p),
        ("d_gmap", ctypes.c_void_p),
        ("gpitch", ctypes.c_int),
    ]


class toed_tune_t(ctypes.Structure):
    _fields_ = [
        ("tile_rows", ctypes.c_int),
        ("persistent", ctypes.c_int),
        ("wg_per_cu", ctypes.c_int),
        ("ms", ctypes.c_float),
    ]


# exported symbols and their signatures: (name, restype, argtypes)
_VP, _I, _F, _LL = ctypes.c_void_p, ctypes.c_int, ctypes.c_float, ctypes.c_longlong
_PI, _PF = ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_float)
_U, _PU = ctypes.c_uint, ctypes.POINTER(ctypes.c_uint)
SIGNATURES = [
    ("toed_version", _I, []),
    ("toed_sigma_radius", _I, [_F, _PI, _PI]),
    ("toed_create", _I, [_I, ctypes.POINTER(_VP)]),
    ("toed_destroy", _I, [_VP]),
    ("toed_set_stream", _I, [_VP, _VP]),
    ("toed_get_stream", _VP, [_VP]),
    ("toed_use_own_stream", _I, [_VP]),
    ("toed_get_device", _I, [_VP]),
    ("toed_backward", _I, [_VP, _I, ctypes.POINTER(toed_grad_t)]),
    ("toed_batch_backward", _I, [_VP, ctypes.POINTER(toed_grad_out_t)]),
    ("toed_set_tile_rows", _I, [_VP, _I]),
    ("toed_set_staging_fraction", _I, [_VP, _F]),
    ("toed_set_orientation_mode", _I, [_VP, _I]),
    ("toed_last_times", _I, [_VP, _PF, _PF]),
    ("toed_detect", _I, [_VP, ctypes.POINTER(toed_image_t), _VP, _VP]),
    ("toed_detect_sync", _I, [_VP, ctypes.POINTER(toed_image_t), ctypes.POINTER(_LL)]),
    ("toed_detect_noalloc", _I, [_VP, ctypes.POINTER(toed_image_t), _VP, _VP]),
    ("toed_batch_plan", _I, [_VP, _I, ctypes.POINTER(toed_image_t), ctypes.POINTER(_VP)]),
    ("toed_batch_set_persistent", _I, [_VP, _I]),
    ("toed_batch_run", _I, [_VP, _VP, _VP]),
    ("toed_batch_graph_run", _I, [_VP, _VP, _VP]),
    ("toed_batch_destroy", _I, [_VP]),
    ("toed_batch_set_grid", _I, [_VP, _I]),
    ("toed_batch_set_images", _I, [_VP, ctypes.POINTER(_VP)]),
    ("toed_batch_set_compact_output", _I, [_VP, _VP, _VP, _LL]),
    ("toed_image_load", _I, [ctypes.c_char_p, _I, ctypes.POINTER(ctypes.POINTER(ctypes.c_ubyte)), _PI, _PI]),
    ("toed_image_free", None, [ctypes.POINTER(ctypes.c_ubyte)]),
    ("toed_batch_autotune", _I, [_VP, _I, ctypes.POINTER(toed_image_t), _I, ctypes.POINTER(toed_tune_t),
                                 ctypes.POINTER(toed_tune_t), _I, _PI]),
    ("toed_subpix_radius", _I, [ctypes.c_double, _PI]),
    ("toed_subpix_plan", _I, [_VP, _I, ctypes.POINTER(toed_subpix_map_t), ctypes.c_double, ctypes.c_double,
                              ctypes.POINTER(_VP)]),
    ("toed_subpix_plan_n", _I, [_VP, _I, ctypes.POINTER(toed_subpix_map_t), ctypes.c_double, ctypes.c_double, _I,
                                ctypes.POINTER(_VP)]),
    ("toed_subpix_run", _I, [_VP, _VP, _VP]),
    ("toed_subpix_destroy", _I, [_VP]),
    ("toed_subpix_set_maps", _I, [_VP, ctypes.POINTER(_VP)]),
    ("toed_subpix_set_timing", _I, [_VP, _I]),
    ("toed_subpix_last_times", _I, [_VP, _PF, _PF, _PF]),
    ("toed_subpix_set_grad", _I, [_VP, _I]),
    ("toed_subpix_backward", _I, [_VP, ctypes.POINTER(toed_subpix_grad_t)]),
    ("toed_gpu_convolve_f32", _I, [_I, _I, _I, _I, _I] + [_VP] * 13),
    ("toed_gpu_nms_f32", _I, [_I, _I, _I, _VP, _VP, _VP, _VP, _VP]),
    ("toed_gpu_convolve_f64", _I, [_I, _I, _I, _I, _I] + [_VP] * 13),
    ("toed_gpu_nms_f64", _I, [_I, _I, _I, _VP, _VP, _VP, _VP, _VP]),
    ("toed_gpu_convolve_sigma_f32", _I, [_I, _I, _I, _I, _I, _VP, _VP, _VP, _VP, _VP, _F]),
    ("toed_gpu_nms_thresh_f32", _I, [_I, _I, _I, _VP, _VP, _VP, _VP, _VP, _F, _I]),
    ("toed_gpu_convolve_sigma_f64", _I, [_I, _I, _I, _I, _I, _VP, _VP, _VP, _VP, _VP, ctypes.c_double]),
    ("toed_gpu_nms_thresh_f64", _I, [_I, _I, _I, _VP, _VP, _VP, _VP, _VP, ctypes.c_double, _I]),
    ("toed_curvelet_create", _I, [_VP, _U, _U, _U] + [ctypes.c_double] * 6 + [_U, _I, _I, ctypes.POINTER(_VP)]),
    ("toed_curvelet_execute", _I, [_VP, _I]),
    ("toed_curvelet_output_size", _I, [_VP, _U, _PU, _PU, _PU]),
    ("toed_curvelet_output", _I, [_VP, _U, _VP, _VP]),
    ("toed_curvelet_destroy", _I, [_VP]),
]

_LIB = None


def lib():
    """Load libtoed_hip.so (raises if it has not been built)."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libtoed_hip.so is not built (run __graft_entry__.build()); "
                               "there is no fallback compute path")
        L = ctypes.CDLL(LIB_PATH)
        for name, res, args in SIGNATURES:
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _LIB = L
    return _LIB


def check(code, what=""):
    if code != 0:
        raise ToedError(code, what)
    return code
