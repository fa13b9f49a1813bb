"""Differentiable TOED: a torch.autograd.Function over the C ABI.

    edges, offsets = detect_differentiable(images, sigma)

``images`` is a float32 [B, H, W] device tensor, ``sigma`` a float or a 0-dim
tensor (which may require grad; on the images' device, its gradient stays
there, with no synchronisation for it, and its value is read on the host once
per in-place update).  The forward is the batched FP32 detection
(one plan over the B images); ``edges`` is the concatenation of the B
raster-ordered N_b x 4 lists and ``offsets`` (int64, B+1, on the host) marks
where each image's rows start.  The backward holds the edgel set fixed (each
edgel keeps its interp pixel, and its octant is re-decided from the FP64
gradient, which only differs at exact ties) and back-propagates
dL/d[x, y, orientation, strength] to every pixel and to sigma (window radii
held fixed) through toed_backward: FP64 recomputation and hand-written adjoint
per edgel, transposed filters scattered into dL/dimage.
"""
import collections
import ctypes

import numpy as np

from . import _lib
from ._lib import check, lib


def _torch():
    import torch
    return torch


class _Plan:
    """A cached forward plan for one (shape, strides, sigma, thresh, border,
    capacity): its images are re-pointed at each forward's tensor
    (toed_batch_set_images), and its rows go to one compact buffer whose
    per-image starts are computed on the device."""

    def __init__(self, det, images, sigma, thresh, border, cap):
        from . import BatchDetector
        B = images.shape[0]
        self.gen = 0    # forward runs of this plan; the tile backward needs the latest
        self.ptr = images.data_ptr()
        self.bd = BatchDetector(det, [images[b] for b in range(B)], [sigma] * B, thresh, border, with_index=True,
                                compact_capacity=B * cap)

    def point_at(self, images):
        if images.data_ptr() != self.ptr:
            self.bd.set_images([images[b] for b in range(images.shape[0])])
            self.ptr = images.data_ptr()
        self.bd.images = []     # no reference to the inputs outlives the call


_PLANS = collections.OrderedDict()
_MAX_PLANS = 2

# sigma held in device memory: its value as last read on the host is kept on
# the tensor object itself, with the storage address and version counter it
# was read at.  An unchanged sigma is then not read again, so a forward does
# not wait for the device's queue (e.g. the previous step's backward) to drain;
# any in-place update bumps the version and is read anew.  (A dictionary keyed
# by tensors would compare them with ==, an elementwise device operation.)
def _sigma_value(sigma_t):
    torch = _torch()
    if not torch.is_tensor(sigma_t) or sigma_t.device.type == "cpu":
        return float(sigma_t.detach()) if torch.is_tensor(sigma_t) else float(sigma_t)
    key = (sigma_t.data_ptr(), sigma_t._version)
    seen = getattr(sigma_t, "_toed_sigma_seen", None)
    if seen is not None and seen[0] == key:
        return seen[1]
    value = float(sigma_t.detach())
    sigma_t._toed_sigma_seen = (key, value)
    return value


def _plan_for(det, images, sigma, thresh, border, cap):
    key = (id(det), str(images.device), tuple(images.shape), tuple(images.stride()), float(sigma), float(thresh),
           int(border), int(cap))
    p = _PLANS.pop(key, None)
    if p is None:
        p = _Plan(det, images, sigma, thresh, border, cap)
    _PLANS[key] = p                   # most recently used last
    while len(_PLANS) > _MAX_PLANS:   # a forward still waiting for its backward keeps its plan alive
        _PLANS.popitem(last=False)
    return p


def _make_function():
    torch = _torch()

    class _Detect(torch.autograd.Function):
        @staticmethod
        def forward(ctx, images, sigma_t, det, thresh, border, cap, sigma):
            plan = _plan_for(det, images, sigma, thresh, border, cap)
            plan.point_at(images)
            bd = plan.bd
            bd.run()
            plan.gen += 1
            ctx.gen = plan.gen
            # the backward's zeroed gradient buffers, enqueued while the forward's
            # kernels run (their host work then costs nothing on the way to the
            # backward launch); only when a backward can follow
            ctx.gbuf = None
            if torch.is_grad_enabled() and (images.requires_grad or (torch.is_tensor(sigma_t) and sigma_t.requires_grad)):
                ctx.gbuf = (torch.zeros(images.shape, dtype=torch.float32, device=images.device),
                            torch.zeros(1, dtype=torch.float64, device=images.device))
            counts, _ = bd.counts_flags()          # the one host read of the forward
            counts = counts.astype(np.int64)
            offsets = np.concatenate([[0], np.cumsum(counts)])
            n = int(offsets[-1])
            # copies, so a later forward on the same plan cannot change this output or backward
            out = bd.cedges[:n].clone()
            ctx.index = bd.cindex[:n].clone()
            ctx.counts_dev = bd.counts.clone()
            ctx.save_for_backward(images)
            ctx.plan = plan
            ctx.counts = counts
            ctx.offsets = offsets
            ctx.sigma = sigma
            ctx.det = det
            ctx.sigma_shape = sigma_t.shape if torch.is_tensor(sigma_t) else None
            ctx.sigma_dtype = sigma_t.dtype if torch.is_tensor(sigma_t) else None
            off_t = torch.from_numpy(offsets)
            ctx.mark_non_differentiable(off_t)
            return out, off_t

        @staticmethod
        def backward(ctx, g_out, g_offsets):
            (images,) = ctx.saved_tensors     # raises if the input was modified in place
            plan, counts, offsets = ctx.plan, ctx.counts, ctx.offsets
            B, H, W = images.shape
            g_out = g_out.contiguous().float()
            if ctx.gbuf is not None:
                gimg, gsig = ctx.gbuf
                ctx.gbuf = None
            else:
                gimg = torch.zeros((B, H, W), dtype=torch.float32, device=images.device)
                gsig = torch.zeros(1, dtype=torch.float64, device=images.device)
            if plan.gen == ctx.gen and plan.ptr == images.data_ptr():
                # the plan's staging still holds this forward's edgels: tile-driven backward,
                # per-image addresses computed from the two batch tensors (an image with
                # no edgels reads no upstream row)
                if g_out.numel() == 0:
                    g_out = torch.zeros((1, 4), dtype=torch.float32, device=images.device)
                g0, i0 = g_out.data_ptr(), gimg.data_ptr()
                gout_ptrs = np.where(counts > 0, g0 + 16 * offsets[:-1], g0).astype(np.uint64)
                gimg_ptrs = i0 + 4 * H * W * np.arange(B, dtype=np.uint64)
                plan.bd.backward_ptrs(gout_ptrs, gimg_ptrs, W, gsig.data_ptr())
                gs = None
                if ctx.sigma_shape is not None:
                    gs = gsig.reshape(()).to(ctx.sigma_dtype).reshape(ctx.sigma_shape)
                return gimg, gs, None, None, None, None, None
            arr = (_lib.toed_grad_t * B)()
            for b in range(B):
                it = arr[b]
                it.in_type = _lib.TOED_F32
                it.d_img = images[b].data_ptr()
                it.H, it.W, it.pitch = H, W, int(images.stride(1))
                it.sigma = ctx.sigma
                it.d_index = ctx.index[int(offsets[b]):].data_ptr() if counts[b] else ctx.counts_dev.data_ptr()
                it.d_count = ctx.counts_dev[b:].data_ptr()
                it.max_edgels = int(counts[b])
                it.d_gout = g_out[int(offsets[b]):].data_ptr() if counts[b] else g_out.data_ptr()
                it.d_gimg = gimg[b].data_ptr()
                it.gpitch = W
                it.d_gsigma = gsig.data_ptr()
            ctx.det._bind_stream()
            check(lib().toed_backward(ctx.det._ctx, B, arr), "toed_backward")
            gs = None
            if ctx.sigma_shape is not None:
                gs = gsig.reshape(()).to(ctx.sigma_dtype).reshape(ctx.sigma_shape)
            return gimg, gs, None, None, None, None, None

    return _Detect


_FN = None


def detect_differentiable(images, sigma, detector=None, thresh=2.0, border=10, capacity=None):
    """Differentiable batched detection; see the module docstring."""
    global _FN
    torch = _torch()
    if _FN is None:
        _FN = _make_function()
    if images.dim() == 2:
        images = images.unsqueeze(0)
    if images.dtype != torch.float32 or not images.is_cuda or images.dim() != 3:
        raise TypeError("images must be a float32 [B, H, W] device tensor")
    if images.stride(2) != 1:
        images = images.contiguous()
    det = detector or _default_detector(images.device)
    B, H, W = images.shape
    cap = int(capacity) if capacity else max(1024, H * W)
    sig = sigma if torch.is_tensor(sigma) else torch.tensor(float(sigma), dtype=torch.float64)
    # sigma's value on the host (read from the user's own tensor object, where
    # a device sigma's last read is remembered)
    return _FN.apply(images, sig, det, float(thresh), int(border), cap, _sigma_value(sig))


_DET = {}


def _default_detector(device):
    from . import Detector
    key = device.index or 0
    if key not in _DET:
        _DET[key] = Detector(key, staging_fraction=0.5)
    return _DET[key]

