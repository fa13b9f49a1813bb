"""MI355X-native TOED (third-order edge detector) detection path.

Python host mirror over the C ABI of libtoed_hip.so (include/toed_capi.h).
PyTorch is used for device memory and streams only; every computation runs
in the package's own HIP kernels.

    Detector                    one device context (kernels on torch's current stream)
    Detector.detect             single image -> raster-ordered N x 4 edgel list
    BatchDetector               many images of uneven size and sigma, one launch
                                sequence (optionally a persistent tile kernel and a
                                captured hipGraph)
    ThirdOrderEdgeDetectionGPU  mirror of the reference class (gpu_toed.hpp:21-74)
    detect_differentiable       torch.autograd.Function: gradients of the edgel
                                rows w.r.t. the pixels and sigma (edgel set frozen)
    SubpixBatch                 subpixel third-order correction of dense edge maps
    SubpixelTOCorrectionGPU     mirror of SubpixelTOCorrectionCPU<T>
                                (subpix_to_correction.hpp:20-96)
    subpix_differentiable       torch.autograd.Function over the correction path:
                                gradients of the edgel rows w.r.t. the maps
    CurveletFormation           curvelet formation on an edgel list, mirror of
                                form_curvelet_process (host code, bit-identical)
    form_curvelets              mirror of curvelet_formation (form_curvelet_main.cpp)
"""
import ctypes
import os
import weakref

import numpy as np

from . import _lib
from ._lib import FLAG_CAPACITY, FLAG_STAGING, TOED_F32, TOED_U8, ToedError, check, lib

__all__ = ["Detector", "BatchDetector", "ThirdOrderEdgeDetectionGPU", "ToedError", "sigma_radius",
           "detect_differentiable", "load_image", "DEFAULT_SIGMA", "DEFAULT_THRESH", "DEFAULT_BORDER", "SubpixBatch",
           "SubpixelTOCorrectionGPU", "subpix_radius", "subpix_differentiable"]

# reference constants: sigma = 2 / 17-tap window (main_gpu_cpu.cpp:158-159),
# threshold 2 (cpu_toed.cpp:403), NMS border 10 (cpu_toed.cpp:400-401)
DEFAULT_SIGMA = 2.0
DEFAULT_THRESH = 2.0
DEFAULT_BORDER = 10


def _torch():
    import torch
    return torch


def sigma_radius(sigma):
    """(R0, R1): the unshifted phase-0 window radius ceil(4 sigma) and R0 + 1."""
    r0, r1 = ctypes.c_int(), ctypes.c_int()
    check(lib().toed_sigma_radius(float(sigma), ctypes.byref(r0), ctypes.byref(r1)), "toed_sigma_radius")
    return r0.value, r1.value


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def _image_struct(img, sigma, thresh, border, edges, capacity, index=None):
    torch = _torch()
    if img.dim() != 2 or not img.is_cuda:
        raise ValueError("image must be a 2-D CUDA/HIP tensor")
    if img.dtype == torch.uint8:
        in_type = TOED_U8
    elif img.dtype == torch.float32:
        in_type = TOED_F32
    else:
        raise TypeError("image dtype must be uint8 or float32, got %s" % img.dtype)
    if img.stride(1) != 1:
        raise ValueError("image rows must be contiguous")
    s = _lib.toed_image_t()
    s.in_type = in_type
    s.d_img = img.data_ptr()
    s.H, s.W = int(img.shape[0]), int(img.shape[1])
    s.pitch = int(img.stride(0))
    s.sigma, s.thresh, s.border = float(sigma), float(thresh), int(border)
    s.d_edges = edges.data_ptr() if edges is not None else None
    s.capacity = int(capacity)
    s.d_index = index.data_ptr() if index is not None else None
    return s


def load_image(path, pgm_raw_offset=False):
    """Read an 8-bit PGM (P5) or PNG into a uint8 H x W numpy array with the
    library's decoder (``toed_image_load``; host code, no GPU needed).  With
    ``pgm_raw_offset`` a PGM is read as the reference's no-OpenCV path reads
    it (main_gpu_cpu.cpp:131 leaves the byte after the header unread, so
    pixel k is the file's byte k - 1)."""
    px = ctypes.POINTER(ctypes.c_ubyte)()
    H, W = ctypes.c_int(), ctypes.c_int()
    check(lib().toed_image_load(os.fsencode(path), 1 if pgm_raw_offset else 0, ctypes.byref(px), ctypes.byref(H),
                                ctypes.byref(W)), "toed_image_load(%s)" % path)
    try:
        return np.ctypeslib.as_array(px, shape=(H.value, W.value)).copy()
    finally:
        lib().toed_image_free(px)


class Detector:
    """A TOED context on one GPU.  Kernels are enqueued on torch's current
    stream of that device, so they order with surrounding torch work."""

    def __init__(self, device=0, tile_rows=0, staging_fraction=1.0):
        torch = _torch()
        self.device = torch.device("cuda", device) if isinstance(device, int) else torch.device(device)
        self._ctx = ctypes.c_void_p()
        check(lib().toed_create(self.device.index or 0, ctypes.byref(self._ctx)), "toed_create")
        if tile_rows:
            check(lib().toed_set_tile_rows(self._ctx, int(tile_rows)), "toed_set_tile_rows")
        check(lib().toed_set_staging_fraction(self._ctx, float(staging_fraction)), "toed_set_staging_fraction")
        self._batches = weakref.WeakSet()   # plans made on this context, destroyed before it

    def close(self):
        if getattr(self, "_ctx", None):
            for b in list(self._batches):
                b.close()
            lib().toed_destroy(self._ctx)
            self._ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _bind_stream(self):
        torch = _torch()
        s = torch.cuda.current_stream(self.device).cuda_stream
        check(lib().toed_set_stream(self._ctx, ctypes.c_void_p(s)), "toed_set_stream")

    def set_tile_rows(self, tile_rows):
        check(lib().toed_set_tile_rows(self._ctx, int(tile_rows)), "toed_set_tile_rows")

    def set_orientation_mode(self, mode):
        """0: third-order orientation at the edgel's interp pixel (cpu_toed.cpp:
        227-232, the default); 1: at its subpixel location, the responses
        interpolated by cubic convolution from the 4 x 4 interp pixels around
        it (the MATLAB detector's QaD path).  Applies to later detections and
        plans."""
        check(lib().toed_set_orientation_mode(self._ctx, int(mode)), "toed_set_orientation_mode")

    def autotune_batch(self, images, sigmas, thresh=DEFAULT_THRESH, border=DEFAULT_BORDER, capacities=None,
                       reps=5, max_results=32):
        """Time the batched form's launch configurations on these images
        (``toed_batch_autotune``: tile rows 8 / 16, one workgroup per tile or
        a persistent grid of 1 .. what fits workgroups per CU) and keep the
        fastest tile rows on this context.  Returns ``{"best": cfg, "results":
        [cfg, ...]}``, cfg = dict(tile_rows, persistent, wg_per_cu, ms); pass
        ``persistent`` and ``wg_per_cu`` of the best to :class:`BatchDetector`."""
        torch = _torch()
        images = list(images)
        n = len(images)
        if n == 0:
            raise ValueError("empty batch")
        sig = list(sigmas) if np.ndim(sigmas) else [float(sigmas)] * n
        caps = [int(c) for c in (capacities or [4 * int(im.shape[0]) * int(im.shape[1]) for im in images])]
        offs = np.concatenate([[0], np.cumsum(caps)]).astype(np.int64)
        edges = torch.empty((int(offs[-1]), 4), dtype=torch.float32, device=images[0].device)
        arr = (_lib.toed_image_t * n)()
        for m, im in enumerate(images):
            arr[m] = _image_struct(im, sig[m], thresh, border, edges[int(offs[m]):], caps[m])
        best = _lib.toed_tune_t()
        res = (_lib.toed_tune_t * max_results)()
        nres = ctypes.c_int(0)
        self._bind_stream()
        check(lib().toed_batch_autotune(self._ctx, n, arr, int(reps), ctypes.byref(best), res, int(max_results),
                                        ctypes.byref(nres)), "toed_batch_autotune")

        def as_dict(t):
            return dict(tile_rows=t.tile_rows, persistent=bool(t.persistent), wg_per_cu=t.wg_per_cu, ms=float(t.ms))

        return {"best": as_dict(best), "results": [as_dict(res[k]) for k in range(nres.value)]}

    def last_times(self):
        a, b = ctypes.c_float(), ctypes.c_float()
        check(lib().toed_last_times(self._ctx, ctypes.byref(a), ctypes.byref(b)), "toed_last_times")
        return a.value, b.value

    def detect_into(self, img, edges, counts, sigma=DEFAULT_SIGMA, thresh=DEFAULT_THRESH, border=DEFAULT_BORDER,
                    index=None):
        """Asynchronous detection into preallocated device buffers:
        edges (capacity x 4 float32), counts (>= 2 int32: N, flags), index
        (optional capacity int32 interp indices)."""
        s = _image_struct(img, sigma, thresh, border, edges, edges.shape[0], index)
        self._bind_stream()
        check(lib().toed_detect(self._ctx, ctypes.byref(s), _ptr(counts[0:1]), _ptr(counts[1:2])), "toed_detect")

    def detect_sync_into(self, img, edges, sigma=DEFAULT_SIGMA, thresh=DEFAULT_THRESH, border=DEFAULT_BORDER,
                         index=None):
        """Synchronous detection with kernel timing (toed_detect_sync).
        Returns N; raises ToedError on staging or capacity overflow."""
        s = _image_struct(img, sigma, thresh, border, edges, edges.shape[0], index)
        self._bind_stream()
        n = ctypes.c_longlong()
        check(lib().toed_detect_sync(self._ctx, ctypes.byref(s), ctypes.byref(n)), "toed_detect_sync")
        return n.value

    def detect(self, img, sigma=DEFAULT_SIGMA, thresh=DEFAULT_THRESH, border=DEFAULT_BORDER, capacity=None,
               with_index=False):
        """Detect edgels in one image (2-D uint8 or float32 device tensor).
        Returns the N x 4 float32 device tensor [x, y, orientation, strength]
        (and the int32 interp indices when with_index)."""
        torch = _torch()
        H, W = int(img.shape[0]), int(img.shape[1])
        cap = 4 * H * W if capacity is None else int(capacity)
        edges = torch.empty((max(cap, 1), 4), dtype=torch.float32, device=img.device)
        index = torch.empty(max(cap, 1), dtype=torch.int32, device=img.device) if with_index else None
        counts = torch.zeros(2, dtype=torch.int32, device=img.device)
        self.detect_into(img, edges[:cap] if cap else edges, counts, sigma, thresh, border, index)
        n, flags = (int(v) for v in counts.cpu())
        if flags & FLAG_STAGING:
            raise ToedError(6, "toed_detect (raise staging_fraction)")
        if flags & FLAG_CAPACITY:
            raise ToedError(5, "toed_detect (%d edgels > capacity %d)" % (n, cap))
        if with_index:
            return edges[:n], index[:n]
        return edges[:n]


class BatchDetector:
    """Images of uneven sizes and sigma processed by one plan: one launch per
    kernel over all tiles of all images (or one persistent tile kernel),
    optionally captured once in a hipGraph and replayed.

    Outputs: one device buffer of sum(capacity) x 4 floats; image m's edgels
    start at row ``offsets[m]``; ``counts[m]`` is its edgel count."""

    def __init__(self, detector, images, sigmas, thresh=DEFAULT_THRESH, border=DEFAULT_BORDER, capacities=None,
                 persistent=False, with_index=False, wg_per_cu=0, compact_capacity=None):
        torch = _torch()
        self.det = detector
        self.images = list(images)
        n = self.n = len(self.images)
        if n == 0:
            raise ValueError("empty batch")
        sig = list(sigmas) if np.ndim(sigmas) else [float(sigmas)] * n
        self.compact = compact_capacity is not None
        if self.compact:
            caps = [1] * n      # the per-image buffers are not used with compact output
        else:
            caps = capacities or [4 * int(im.shape[0]) * int(im.shape[1]) for im in self.images]
        self.capacities = [int(c) for c in caps]
        self.offsets = np.concatenate([[0], np.cumsum(self.capacities)]).astype(np.int64)
        dev = self.images[0].device
        self.edges = torch.empty((int(self.offsets[-1]), 4), dtype=torch.float32, device=dev)
        self.index = torch.empty(int(self.offsets[-1]), dtype=torch.int32, device=dev) if with_index else None
        # counts and flags side by side: one device-to-host copy reads both
        self.cf = torch.zeros(2 * n, dtype=torch.int32, device=dev)
        self.counts = self.cf[:n]
        self.flags = self.cf[n:]
        arr = (_lib.toed_image_t * n)()
        for m, im in enumerate(self.images):
            o = int(self.offsets[m])
            arr[m] = _image_struct(im, sig[m], thresh, border, self.edges[o:], self.capacities[m],
                                   self.index[o:] if with_index else None)
        self._arr = arr
        self._b = ctypes.c_void_p()
        detector._bind_stream()
        check(lib().toed_batch_plan(detector._ctx, n, arr, ctypes.byref(self._b)), "toed_batch_plan")
        detector._batches.add(self)
        if persistent:
            check(lib().toed_batch_set_persistent(self._b, 1), "toed_batch_set_persistent")
            if wg_per_cu:
                check(lib().toed_batch_set_grid(self._b, int(wg_per_cu)), "toed_batch_set_grid")
        if self.compact:
            cap = int(compact_capacity)
            self.cedges = torch.empty((max(cap, 1), 4), dtype=torch.float32, device=dev)
            self.cindex = torch.empty(max(cap, 1), dtype=torch.int32, device=dev) if with_index else None
            check(lib().toed_batch_set_compact_output(self._b, _ptr(self.cedges), _ptr(self.cindex), cap),
                  "toed_batch_set_compact_output")

    def set_images(self, images):
        """Point the plan at other device images of the same shapes, dtypes and
        row strides (toed_batch_set_images): no new plan, no new buffers."""
        images = list(images)
        if len(images) != self.n:
            raise ValueError("one image per planned image")
        for a, b in zip(images, self.images):
            if a.shape != b.shape or a.dtype != b.dtype or a.stride() != b.stride() or a.device != b.device:
                raise ValueError("images must match the planned shapes, dtypes and strides")
        arr = (ctypes.c_void_p * self.n)(*[im.data_ptr() for im in images])
        check(lib().toed_batch_set_images(self._b, arr), "toed_batch_set_images")
        self.images = images

    def run(self, graph=False):
        self.det._bind_stream()
        fn = lib().toed_batch_graph_run if graph else lib().toed_batch_run
        check(fn(self._b, _ptr(self.counts), _ptr(self.flags)), "toed_batch_run")

    def counts_flags(self):
        """Synchronise and return (counts, flags) as numpy arrays (one copy);
        raises on staging or capacity overflow."""
        cf = self.cf.cpu().numpy()
        counts, flags = cf[:self.n], cf[self.n:]
        if (flags & FLAG_STAGING).any():
            raise ToedError(6, "batch (raise staging_fraction)")
        if (flags & FLAG_CAPACITY).any():
            raise ToedError(5, "batch (edgel capacity)")
        return counts, flags

    def _starts(self, counts):
        if self.compact:
            return np.concatenate([[0], np.cumsum(counts)[:-1]]).astype(np.int64)
        return self.offsets[:-1]

    def results(self):
        """Synchronise and return the per-image N x 4 device tensors."""
        counts, _ = self.counts_flags()
        buf = self.cedges if self.compact else self.edges
        st = self._starts(counts)
        return [buf[int(st[m]):int(st[m]) + int(counts[m])] for m in range(self.n)]

    def backward(self, gouts, gimgs, gsigma=None):
        """Adjoint of the last run (edgel set frozen): gouts[m] holds image m's
        dL/d[x, y, orientation, strength] rows in its output layout, gimgs[m]
        receives dL/dimage (accumulated), gsigma (float64 device tensor, one
        element, optional) dL/dsigma summed over the images."""
        n = self.n
        arr = (_lib.toed_grad_out_t * n)()
        for m in range(n):
            arr[m].d_gout = gouts[m].data_ptr()
            arr[m].d_gimg = gimgs[m].data_ptr()
            arr[m].gpitch = int(gimgs[m].stride(0))
            arr[m].d_gsigma = gsigma.data_ptr() if gsigma is not None else None
        self.det._bind_stream()
        check(lib().toed_batch_backward(self._b, arr), "toed_batch_backward")

    def backward_ptrs(self, gout_ptrs, gimg_ptrs, gpitches, gsigma_ptr=None):
        """backward() from raw device addresses (ints): per image the first
        upstream row, the first dL/dimage element and its row pitch in
        elements; gsigma_ptr the one-element float64 accumulator or None.  For
        callers that already hold the batch in single tensors (no per-image
        views)."""
        n = self.n
        arr = getattr(self, "_grad_arr", None)
        if arr is None:
            # toed_grad_out_t as a numpy record array: filled with vector
            # assignments instead of one ctypes field write per image
            dt = np.dtype({"names": ["d_gout", "d_gimg", "gpitch", "d_gsigma"],
                           "formats": [np.uint64, np.uint64, np.int32, np.uint64],
                           "offsets": [_lib.toed_grad_out_t.d_gout.offset, _lib.toed_grad_out_t.d_gimg.offset,
                                       _lib.toed_grad_out_t.gpitch.offset, _lib.toed_grad_out_t.d_gsigma.offset],
                           "itemsize": ctypes.sizeof(_lib.toed_grad_out_t)})
            arr = self._grad_arr = np.zeros(n, dt)
        arr["d_gout"] = gout_ptrs
        arr["d_gimg"] = gimg_ptrs
        arr["gpitch"] = gpitches
        arr["d_gsigma"] = gsigma_ptr or 0
        self.det._bind_stream()
        check(lib().toed_batch_backward(self._b, arr.ctypes.data_as(ctypes.POINTER(_lib.toed_grad_out_t))),
              "toed_batch_backward")

    def index_results(self):
        counts = self.counts.cpu().numpy()
        buf = self.cindex if self.compact else self.index
        st = self._starts(counts)
        return [buf[int(st[m]):int(st[m]) + int(counts[m])] for m in range(len(counts))]

    def close(self):
        if getattr(self, "_b", None):
            lib().toed_batch_destroy(self._b)
            self._b = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class ThirdOrderEdgeDetectionGPU:
    """Mirror of the reference's ThirdOrderEdgeDetectionGPU<T>
    (gpu_toed.hpp:21-74): same constructor order (device, H, W, kernel_size,
    sigma), same methods and public fields.

    The fused pipeline computes NMS inside the tile kernel, so
    ``convolve_img`` launches the whole detection and records
    ``time_conv`` (tile kernel) and ``time_nms`` (scan + finalize), both in ms
    like the reference's CUDA-event timings; ``non_maximum_suppresion``
    gathers the list, fills ``TOED_edges`` and writes
    ``data_final_output_gpu.txt`` (gpu_toed.hpp:415-441)."""

    num_of_edge_data = 4

    def __init__(self, device, H, W, kernel_size=17, sigma=2, dtype="float32"):
        # T = float or double: one computation (FP32 dense filters, FP64 per
        # edgel); for double the FP32 rows are widened (main_gpu_cpu.cpp:200-209)
        if np.dtype(dtype) not in (np.dtype(np.float32), np.dtype(np.float64)):
            raise NotImplementedError("T is float32 or float64")
        self.dtype = np.dtype(dtype)
        sig = float(sigma)
        r0, _ = sigma_radius(sig)
        if int(kernel_size) != 2 * r0 + 1:
            raise ValueError("kernel_size %d does not match sigma %g (expected %d)" % (kernel_size, sig, 2 * r0 + 1))
        torch = _torch()
        self.device_id = int(device)
        self.img_height, self.img_width = int(H), int(W)
        self.interp_img_height, self.interp_img_width = 2 * int(H), 2 * int(W)
        self.kernel_sz, self.shifted_kernel_sz, self.g_sig = int(kernel_size), int(kernel_size) + 2, sig
        self.output_dir = "./output_files"
        self.edge_pt_list_idx = 0
        self.time_conv = 0.0
        self.time_nms = 0.0
        self._det = Detector(self.device_id)
        dev = self._det.device
        self._img = torch.zeros((self.img_height, self.img_width), dtype=torch.uint8, device=dev)
        cap = self.interp_img_height * self.interp_img_width     # main_gpu_cpu.cpp:56-66
        self._edges = torch.empty((cap, 4), dtype=torch.float32, device=dev)
        self.subpix_edge_pts_final = np.zeros((0, 4), self.dtype)
        self._n = 0

    def preprocessing(self, image):
        """Load the H x W grey image: a uint8 array (the cv::Mat overload,
        gpu_toed.hpp:140-184) or a binary stream of H*W bytes (the ifstream
        overload, gpu_toed.hpp:189-231)."""
        torch = _torch()
        if hasattr(image, "read"):
            raw = image.read(self.img_height * self.img_width)
            arr = np.frombuffer(raw, np.uint8)
            if arr.size != self.img_height * self.img_width:
                raise ValueError("stream ended after %d of %d bytes" % (arr.size, self.img_height * self.img_width))
            arr = arr.reshape(self.img_height, self.img_width)
        else:
            arr = np.ascontiguousarray(image)
            if arr.shape != (self.img_height, self.img_width):
                raise ValueError("image shape %s != (%d, %d)" % (arr.shape, self.img_height, self.img_width))
            arr = arr.astype(np.uint8, copy=False)
        self._img.copy_(torch.from_numpy(arr))

    def convolve_img(self):
        self._n = self._det.detect_sync_into(self._img, self._edges, self.g_sig, DEFAULT_THRESH, DEFAULT_BORDER)
        self.time_conv, self.time_nms = self._det.last_times()

    def non_maximum_suppresion(self, TOED_edges=None):
        n = self._n
        self.subpix_edge_pts_final = self._edges[:n].cpu().numpy().astype(self.dtype, copy=False)
        if TOED_edges is not None:
            TOED_edges[:n, :4] = self.subpix_edge_pts_final
        self.edge_pt_list_idx = n
        self.write_array_to_file("data_final_output_gpu.txt", self.subpix_edge_pts_final, n, 4)
        return n

    def set_output_dir(self, d):
        self.output_dir = d if d else "./output_files"

    def write_array_to_file(self, filename, wr_data, first_dim, second_dim):
        """Tab-separated text, precision 10 (gpu_toed.hpp:452-478)."""
        path = os.path.join(self.output_dir, filename)
        data = np.asarray(wr_data).reshape(-1)[: first_dim * second_dim].reshape(first_dim, second_dim)
        try:
            with open(path, "w") as f:
                for row in data:
                    f.write("".join("%.10g\t" % v for v in row) + "\n")
        except OSError:
            print("write data file cannot be opened!")

    def read_array_from_file(self, filename, rd_data, first_dim, second_dim):
        """Reads whitespace-separated values from ./test_files/<filename>
        (gpu_toed.hpp:483-509) into rd_data (flat, row-major)."""
        path = os.path.join("./test_files", filename)
        if not os.path.exists(path):
            print("input read file not existed!")
            return
        vals = np.loadtxt(path, dtype=np.float64).reshape(-1)
        flat = rd_data.reshape(-1)
        k = min(flat.size, vals.size, first_dim * second_dim)
        flat[:k] = vals[:k]


def detect_differentiable(images, sigma, detector=None, thresh=DEFAULT_THRESH, border=DEFAULT_BORDER, capacity=None):
    """Differentiable batched detection: returns (edges [N, 4], offsets [B+1]);
    see the autograd module."""
    from .autograd import detect_differentiable as _dd
    return _dd(images, sigma, detector, thresh, border, capacity)


from .subpix import SubpixBatch, SubpixelTOCorrectionGPU, subpix_differentiable, subpix_radius  # noqa: E402
from .curvelet import CurveletFormation, form_curvelets  # noqa: E402
