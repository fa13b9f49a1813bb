"""Deterministic synthetic inputs (ctypes over libtoed_scene.so).

The generators are C (csrc/toed_scene.c) with an integer SplitMix64 RNG, so
this container and any GPU host produce bit-identical images from a seed.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


def _lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(HERE, "libtoed_scene.so")
        if not os.path.exists(path):
            raise RuntimeError("libtoed_scene.so is not built; run __graft_entry__.build()")
        L = ctypes.CDLL(path)
        L.toed_scene_u8.argtypes = [ctypes.c_uint64, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                    ctypes.c_double, ctypes.c_void_p]
        L.toed_scene_u8.restype = ctypes.c_int
        L.toed_bezier_map_f32.argtypes = [ctypes.c_uint64, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                          ctypes.c_void_p]
        L.toed_bezier_map_f32.restype = ctypes.c_int
        L.toed_loss_weights.argtypes = [ctypes.c_uint64, ctypes.c_void_p]
        L.toed_scene_default_shapes.argtypes = [ctypes.c_int, ctypes.c_int]
        _LIB = L
    return _LIB


def shapes_scene(seed, H, W, n_shapes=0, size_scale=0.0):
    """uint8 H x W scene: background ~U[40,200], 24*k disks/convex polygons
    (k = max(1, H // 540)), 4x4 supersampled, N(0, 2^2) noise."""
    out = np.empty((H, W), np.uint8)
    if _lib().toed_scene_u8(int(seed) & (2**64 - 1), int(H), int(W), int(n_shapes), float(size_scale),
                            out.ctypes.data):
        raise ValueError("scene generation failed")
    return out


def bezier_map(seed, H, W, n_curves=40):
    """float32 H x W edge map from random cubic Bezier curves."""
    out = np.empty((H, W), np.float32)
    if _lib().toed_bezier_map_f32(int(seed) & (2**64 - 1), int(H), int(W), int(n_curves), out.ctypes.data):
        raise ValueError("bezier map generation failed")
    return out


def loss_weights(seed):
    """(a, b, c, d) ~ N(0, 1) for the edgel loss sum(a x + b y + c theta + d s)."""
    w = np.empty(4, np.float64)
    _lib().toed_loss_weights(int(seed) & (2**64 - 1), w.ctypes.data)
    return w
