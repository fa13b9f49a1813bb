"""Subpixel third-order correction of dense edge maps (GPU form of the
reference's subpix_TO_correction/ path) over the C ABI of libtoed_hip.so.

    SubpixBatch               many float32 maps of any sizes, one plan and one
                              launch sequence -> per-map raster-ordered N x 4
                              lists [x, y, orientation, strength], optional
                              TO_edgemap / TO_orientation maps
    SubpixelTOCorrectionGPU   mirror of SubpixelTOCorrectionCPU<T>
                              (subpix_to_correction.hpp:20-96)
"""
import collections
import ctypes
import os

import numpy as np

from . import _lib
from ._lib import FLAG_CAPACITY, ToedError, check, lib

# reference defaults (main_subpix_to_correction.cpp:147-151)
DEFAULT_THRESH = 0.2
DEFAULT_SIGMA = 2.0
DEFAULT_NBR_RADIUS = 2.0


def _torch():
    import torch
    return torch


def subpix_radius(sigma):
    """Filter radius ceil(4 sigma) of the correction path (subpix_to_correction.cpp:107)."""
    r = ctypes.c_int()
    check(lib().toed_subpix_radius(float(sigma), ctypes.byref(r)), "toed_subpix_radius")
    return r.value


def work_shape(H, W, interp_n=0):
    """Work grid of an H x W map with interp_n > 0 (subpix_to_correction.cpp:102-104)."""
    s = 1 << int(interp_n)
    return (H - 1) * s + 1, (W - 1) * s + 1


class SubpixBatch:
    """Edge maps (2-D float32 device tensors, row stride >= width, unit column
    stride) corrected by one plan.  ``orients`` (optional, same shapes, radians)
    gives the orientation of every map, or None per map to estimate it from
    the map.  ``interp_n`` > 0 runs every map on its bilinearly upsampled work
    grid ((H - 1) 2^n + 1) x ((W - 1) 2^n + 1) (toed_subpix_plan_n); x, y stay
    in map pixels and ``index`` holds work-grid pixels.  Outputs: one device
    buffer of sum(capacity) x 4 floats (capacity defaults to the work grid's
    pixel count); map m's edgels start at row ``offsets[m]``, ``counts[m]`` of
    them."""

    def __init__(self, detector, maps, orients=None, thresh=DEFAULT_THRESH, sigma=DEFAULT_SIGMA,
                 nbr_radius=DEFAULT_NBR_RADIUS, capacities=None, with_index=False, with_maps=False, interp_n=0):
        torch = _torch()
        self.interp_n = int(interp_n)
        self.det = detector
        self.maps = list(maps)
        n = len(self.maps)
        if n == 0:
            raise ValueError("empty batch")
        ors = list(orients) if orients is not None else [None] * n
        if len(ors) != n:
            raise ValueError("one orientation entry per map")
        thr = list(thresh) if np.ndim(thresh) else [float(thresh)] * n
        for m, E in enumerate(self.maps):
            if E.dim() != 2 or not E.is_cuda or E.dtype != torch.float32 or E.stride(1) != 1:
                raise TypeError("maps must be 2-D float32 device tensors with unit column stride")
            if ors[m] is not None and (ors[m].shape != E.shape or ors[m].dtype != torch.float32
                                       or ors[m].stride(1) != 1):
                raise TypeError("orientation maps must match their edge maps (float32)")
        caps = capacities or [int(np.prod(work_shape(int(E.shape[0]), int(E.shape[1]), self.interp_n)))
                              for E in self.maps]
        self.capacities = [int(c) for c in caps]
        self.offsets = np.concatenate([[0], np.cumsum(self.capacities)]).astype(np.int64)
        dev = self.maps[0].device
        self.edges = torch.empty((max(1, int(self.offsets[-1])), 4), dtype=torch.float32, device=dev)
        self.index = torch.empty(max(1, int(self.offsets[-1])), dtype=torch.int32, device=dev) if with_index else None
        self.counts = torch.zeros(n, dtype=torch.int32, device=dev)
        self.flags = torch.zeros(n, dtype=torch.int32, device=dev)
        self.to_edgemap = [torch.empty_like(E, memory_format=torch.contiguous_format) for E in self.maps] \
            if with_maps else None
        self.to_orientation = [torch.empty_like(E, memory_format=torch.contiguous_format) for E in self.maps] \
            if with_maps else None
        arr = (_lib.toed_subpix_map_t * n)()
        for m, E in enumerate(self.maps):
            it = arr[m]
            o = int(self.offsets[m])
            it.d_map = E.data_ptr()
            it.H, it.W, it.pitch = int(E.shape[0]), int(E.shape[1]), int(E.stride(0))
            if ors[m] is not None:
                it.d_orient = ors[m].data_ptr()
                it.opitch = int(ors[m].stride(0))
            it.thresh = float(thr[m])
            it.d_edges = self.edges[o:].data_ptr()
            it.capacity = self.capacities[m]
            it.d_index = self.index[o:].data_ptr() if with_index else None
            if with_maps:
                it.d_to_edgemap = self.to_edgemap[m].data_ptr()
                it.d_to_orientation = self.to_orientation[m].data_ptr()
                it.tpitch = int(E.shape[1])
        self._orients = ors
        self._arr = arr
        self._p = ctypes.c_void_p()
        detector._bind_stream()
        check(lib().toed_subpix_plan_n(detector._ctx, n, arr, float(sigma), float(nbr_radius), self.interp_n,
                                       ctypes.byref(self._p)), "toed_subpix_plan_n")
        detector._batches.add(self)

    def run(self):
        self.det._bind_stream()
        check(lib().toed_subpix_run(self._p, ctypes.c_void_p(self.counts.data_ptr()),
                                    ctypes.c_void_p(self.flags.data_ptr())), "toed_subpix_run")

    def set_maps(self, maps):
        """Point the plan at other device maps of the same shapes and row
        strides (toed_subpix_set_maps): no new plan, no new buffers."""
        maps = list(maps)
        if len(maps) != len(self.maps):
            raise ValueError("one map per planned map")
        for a, b in zip(maps, self.maps):
            if b is not None and (a.shape != b.shape or a.dtype != b.dtype or a.stride() != b.stride()):
                raise ValueError("maps must match the planned shapes, dtypes and strides")
        arr = (ctypes.c_void_p * len(maps))(*[E.data_ptr() for E in maps])
        check(lib().toed_subpix_set_maps(self._p, arr), "toed_subpix_set_maps")
        self.maps = maps

    def set_timing(self, on=True):
        check(lib().toed_subpix_set_timing(self._p, int(bool(on))), "toed_subpix_set_timing")

    def last_times(self):
        """(orientation + NMS, isolation + convolutions + correction with the
        list rows, rank scans + TO maps) in
        ms for the last run with timing on."""
        a, b, c = ctypes.c_float(), ctypes.c_float(), ctypes.c_float()
        check(lib().toed_subpix_last_times(self._p, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c)),
              "toed_subpix_last_times")
        return a.value, b.value, c.value

    def set_grad(self, on=True):
        """Keep what the adjoint needs on later runs (differentiable form)."""
        check(lib().toed_subpix_set_grad(self._p, int(bool(on))), "toed_subpix_set_grad")

    def backward(self, gouts, gmaps):
        """Adjoint of the last run (made with set_grad on), every discrete
        decision frozen: gouts[m] holds map m's dL/d[x, y, orientation,
        strength] rows in its output layout (float32 device, >= count rows),
        gmaps[m] receives dL/dmap (accumulated, float32 device, unit column
        stride)."""
        n = len(self.maps)
        arr = (_lib.toed_subpix_grad_t * n)()
        for m in range(n):
            arr[m].d_gout = gouts[m].data_ptr()
            arr[m].d_gmap = gmaps[m].data_ptr()
            arr[m].gpitch = int(gmaps[m].stride(0))
        self.det._bind_stream()
        check(lib().toed_subpix_backward(self._p, arr), "toed_subpix_backward")

    def backward_ptrs(self, gout_ptrs, gmap_ptrs, gpitches):
        """backward() from raw device addresses (ints): per map its first
        upstream row, its first dL/dmap element and that plane's row pitch in
        elements.  For callers that hold the batch in single tensors."""
        n = len(self.maps)
        arr = getattr(self, "_grad_arr", None)
        if arr is None:
            arr = self._grad_arr = (_lib.toed_subpix_grad_t * n)()
        for m in range(n):
            it = arr[m]
            it.d_gout = gout_ptrs[m]
            it.d_gmap = gmap_ptrs[m]
            it.gpitch = gpitches[m]
        self.det._bind_stream()
        check(lib().toed_subpix_backward(self._p, arr), "toed_subpix_backward")

    def _counts(self):
        counts = self.counts.cpu().numpy()
        if (self.flags.cpu().numpy() & FLAG_CAPACITY).any():
            raise ToedError(5, "subpix batch (edgel capacity)")
        return counts

    def results(self):
        """Synchronise and return the per-map N x 4 device tensors."""
        counts = self._counts()
        return [self.edges[int(self.offsets[m]):int(self.offsets[m]) + int(counts[m])] for m in range(len(counts))]

    def index_results(self):
        counts = self._counts()
        return [self.index[int(self.offsets[m]):int(self.offsets[m]) + int(counts[m])] for m in range(len(counts))]

    def close(self):
        if getattr(self, "_p", None):
            lib().toed_subpix_destroy(self._p)
            self._p = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class SubpixelTOCorrectionGPU:
    """Mirror of the reference's SubpixelTOCorrectionCPU<T>
    (subpix_to_correction.hpp:20-96, driven as main_subpix_to_correction.cpp:
    279-293): ctor (H, W, threshold, g_sigma, device, n), set_edgemap,
    set_orientation or estimate_orientation, run(edginfo_out) -> N, public
    TO_edgemap / TO_orientation / subpix_edge_pts_final (host arrays),
    set_neighbor_radius, set_output_dir, write_array_to_file.  n = interp_n
    > 0 runs on the bilinearly upsampled work grid (:92-105, :226-240): the
    list holds up to one row per work pixel, TO maps stay H x W."""

    num_of_edge_data = 4

    def __init__(self, H, W, threshold=DEFAULT_THRESH, g_sigma=DEFAULT_SIGMA, device=0, n=0):
        from . import Detector
        torch = _torch()
        self.interp_n = max(0, int(n or 0))        # the reference clamps n < 0 to 0 (:94)
        if self.interp_n > _lib.TOED_SUBPIX_MAX_INTERP:
            raise ValueError("interp_n above %d" % _lib.TOED_SUBPIX_MAX_INTERP)
        self.img_height, self.img_width = int(H), int(W)
        self.work_height, self.work_width = work_shape(self.img_height, self.img_width, self.interp_n)
        self.thresh, self.sigma = float(threshold), float(g_sigma)
        subpix_radius(self.sigma)   # validates sigma
        self.neighbor_radius = DEFAULT_NBR_RADIUS
        self.output_dir = "./output_files"
        self._det = Detector(device)
        self._dev = torch.device("cuda", device)
        self._map = torch.zeros((H, W), dtype=torch.float32, device=self._dev)
        self._orient = None
        self.TO_edgemap = np.zeros((H, W), np.float32)
        self.TO_orientation = np.zeros((H, W), np.float32)
        self.subpix_edge_pts_final = np.zeros((0, 4), np.float32)
        self.edge_pt_list_idx = 0
        self.time_nms = self.time_conv = self.time_correct = 0.0

    def set_edgemap(self, data):
        self._map.copy_(_torch().as_tensor(np.asarray(data, np.float32).reshape(self.img_height, self.img_width)))

    def set_orientation(self, data):
        torch = _torch()
        self._orient = torch.as_tensor(np.asarray(data, np.float32).reshape(self.img_height, self.img_width)).to(
            self._dev)

    def estimate_orientation(self):
        self._orient = None

    def set_neighbor_radius(self, radius):
        self.neighbor_radius = float(radius)

    def set_output_dir(self, d):
        self.output_dir = d or "./output_files"

    def run(self, edginfo_out=None):
        torch = _torch()
        b = SubpixBatch(self._det, [self._map], [self._orient], self.thresh, self.sigma, self.neighbor_radius,
                        with_maps=True, interp_n=self.interp_n)
        try:
            b.set_timing(True)
            b.run()
            edges = b.results()[0].cpu().numpy()
            self.time_nms, self.time_conv, self.time_correct = b.last_times()
            self.TO_edgemap = b.to_edgemap[0].cpu().numpy()
            self.TO_orientation = b.to_orientation[0].cpu().numpy()
        finally:
            b.close()
        self.subpix_edge_pts_final = edges
        self.edge_pt_list_idx = len(edges)
        if edginfo_out is not None:
            edginfo_out.reshape(-1)[:edges.size] = edges.reshape(-1)
        return self.edge_pt_list_idx

    def write_array_to_file(self, filename, wr_data, first_dim, second_dim):
        """Tab-separated text as subpix_to_correction.cpp:741-775."""
        if wr_data is None or first_dim <= 0 or second_dim <= 0:
            return
        os.makedirs(self.output_dir, exist_ok=True)
        a = np.asarray(wr_data).reshape(-1)[:first_dim * second_dim].reshape(first_dim, second_dim)
        with open(os.path.join(self.output_dir, filename), "w") as f:
            for row in a:
                f.write("".join("%.6g\t" % v for v in row) + "\n")

    def close(self):
        self._det.close()


class _SubpixPlan:
    """A cached differentiable plan for one (shape, strides, threshold, sigma,
    radius, capacity), re-pointed at each forward's maps."""

    def __init__(self, det, maps, thresh, sigma, nbr, cap):
        self.gen = 0
        self.ptr = maps.data_ptr()
        self.b = SubpixBatch(det, [maps[k] for k in range(maps.shape[0])], None, thresh, sigma, nbr,
                             capacities=[cap] * maps.shape[0], with_index=True)
        self.b.set_grad(True)

    def point_at(self, maps):
        if maps.data_ptr() != self.ptr:
            self.b.set_maps([maps[k] for k in range(maps.shape[0])])
            self.ptr = maps.data_ptr()
        self.b.maps = [None] * maps.shape[0]    # no reference to the inputs outlives the call


_SP_PLANS = collections.OrderedDict()
_SP_MAX_PLANS = 2


def _subpix_plan(det, maps, thresh, sigma, nbr, cap):
    key = (id(det), str(maps.device), tuple(maps.shape), tuple(maps.stride()), float(thresh), float(sigma),
           float(nbr), int(cap))
    p = _SP_PLANS.pop(key, None)
    if p is None:
        p = _SubpixPlan(det, maps, thresh, sigma, nbr, cap)
    _SP_PLANS[key] = p
    while len(_SP_PLANS) > _SP_MAX_PLANS:
        _SP_PLANS.popitem(last=False)
    return p


_SP_FN = None


def _make_subpix_function():
    torch = _torch()

    class _Subpix(torch.autograd.Function):
        @staticmethod
        def forward(ctx, maps, det, thresh, sigma, nbr, cap):
            plan = _subpix_plan(det, maps, thresh, sigma, nbr, cap)
            plan.point_at(maps)
            b = plan.b
            b.run()
            plan.gen += 1
            ctx.gen = plan.gen
            counts = b._counts().astype(np.int64)
            offsets = np.concatenate([[0], np.cumsum(counts)])
            parts = [b.edges[int(b.offsets[m]):int(b.offsets[m]) + int(counts[m])] for m in range(len(counts))]
            out = torch.cat(parts, 0) if parts else b.edges[:0]
            ctx.save_for_backward(maps)
            ctx.plan, ctx.counts, ctx.offsets = plan, counts, offsets
            off_t = torch.from_numpy(offsets)
            ctx.mark_non_differentiable(off_t)
            return out, off_t

        @staticmethod
        def backward(ctx, g_out, g_offsets):
            (maps,) = ctx.saved_tensors           # raises if the maps were modified in place
            plan, counts, offsets = ctx.plan, ctx.counts, ctx.offsets
            if plan.gen != ctx.gen or plan.ptr != maps.data_ptr():
                # the plan ran again since this forward: run it once more on this
                # forward's maps (the same decisions and records), then take the adjoint
                plan.point_at(maps)
                plan.b.run()
                plan.gen += 1
                assert np.array_equal(plan.b._counts().astype(np.int64), counts)
            B, H, W = maps.shape
            g_out = g_out.contiguous().float()
            if g_out.numel() == 0:
                g_out = torch.zeros((1, 4), dtype=torch.float32, device=maps.device)
            gmap = torch.zeros((B, H, W), dtype=torch.float32, device=maps.device)
            # per-map addresses from the two batch tensors (a map with no edgels reads no row)
            g0, m0 = g_out.data_ptr(), gmap.data_ptr()
            plan.b.backward_ptrs([g0 + 16 * int(offsets[m]) if counts[m] else g0 for m in range(B)],
                                 [m0 + 4 * H * W * m for m in range(B)], [W] * B)
            return gmap, None, None, None, None, None

    return _Subpix


def subpix_differentiable(maps, thresh=DEFAULT_THRESH, sigma=DEFAULT_SIGMA, nbr_radius=DEFAULT_NBR_RADIUS,
                          detector=None, capacity=None):
    """Differentiable subpixel third-order correction of a float32 [B, H, W]
    device tensor of edge maps (orientation estimated from each map).
    Returns (edges, offsets) like detect_differentiable: the concatenated
    raster-ordered N_b x 4 lists and the int64 row offsets (host).  The
    backward freezes every discrete decision of the forward (edgel set, NMS
    octants, cubic stencils, per-pixel half-turn of the direction and
    sign(-Oxy)) and returns dL/dmaps."""
    global _SP_FN
    torch = _torch()
    if _SP_FN is None:
        _SP_FN = _make_subpix_function()
    if maps.dim() == 2:
        maps = maps.unsqueeze(0)
    if maps.dtype != torch.float32 or not maps.is_cuda or maps.dim() != 3:
        raise TypeError("maps must be a float32 [B, H, W] device tensor")
    if maps.stride(2) != 1:
        maps = maps.contiguous()
    from . import autograd as _ag
    det = detector or _ag._default_detector(maps.device)
    B, H, W = maps.shape
    cap = int(capacity) if capacity else max(1024, H * W // 4)
    return _SP_FN.apply(maps, det, float(thresh), float(sigma), float(nbr_radius), cap)
