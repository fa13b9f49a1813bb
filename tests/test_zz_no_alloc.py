"""toed_detect_noalloc, the allocation-free form of toed_detect (scratch from
the library's static per-GPU array), called through the C ABI: two calls in a
row -- the 2160 x 3840 u8 frame at sigma 2, then a smaller f32 image at sigma
1.5 -- each give exactly toed_detect's rows, interp indices and count (same
kernels, same taps), and an image needing more scratch than the array holds
is refused on the host before any launch."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def _call(pkg, fn, ctx, img, sigma, cap):
    edges = torch.full((cap, 4), -7.0, dtype=torch.float32, device="cuda")
    index = torch.full((cap,), -7, dtype=torch.int32, device="cuda")
    cf = torch.full((2,), -7, dtype=torch.int32, device="cuda")
    s = pkg._image_struct(img, sigma, pkg.DEFAULT_THRESH, pkg.DEFAULT_BORDER, edges, cap, index)
    code = fn(ctx, ctypes.byref(s), ctypes.c_void_p(cf[0:1].data_ptr()), ctypes.c_void_p(cf[1:2].data_ptr()))
    torch.cuda.synchronize()
    n, flags = (int(v) for v in cf.cpu())
    return code, n, flags, edges[:max(n, 0)].cpu().numpy(), index[:max(n, 0)].cpu().numpy()


def test_noalloc_matches_detect(pkg, scene):
    L = pkg._lib.lib()
    det = pkg.Detector(0)
    try:
        det._bind_stream()
        big = torch.from_numpy(scene.shapes_scene(2160 * 3840, 2160, 3840)).cuda()
        small = torch.from_numpy(scene.shapes_scene(41, 480, 752).astype(np.float32)).cuda()
        for label, img, sigma in [("2160x3840 u8 sigma 2", big, 2.0), ("480x752 f32 sigma 1.5", small, 1.5)]:
            cap = 4 * img.shape[0] * img.shape[1] // 16
            print("noalloc case:", label, flush=True)
            c0, n0, f0, e0, i0 = _call(pkg, L.toed_detect_noalloc, det._ctx, img, sigma, cap)
            c1, n1, f1, e1, i1 = _call(pkg, L.toed_detect, det._ctx, img, sigma, cap)
            print("  noalloc: status %d, %d edgels, flags %d; toed_detect: status %d, %d edgels, flags %d"
                  % (c0, n0, f0, c1, n1, f1), flush=True)
            assert c0 == 0 and c1 == 0 and f0 == 0 and f1 == 0
            assert n0 == n1 > 1000
            np.testing.assert_array_equal(i0, i1)
            np.testing.assert_array_equal(e0, e1)
        # more scratch than the array holds: refused before anything is launched
        huge = torch.zeros((4096, 4096), dtype=torch.uint8, device="cuda")
        code, _, _, _, _ = _call(pkg, L.toed_detect_noalloc, det._ctx, huge, 2.0, 16)
        assert code == 1   # TOED_ERR_ARG
    finally:
        det.close()
