"""GPU parity of the fused detection path against the FP64 oracle (which is
itself pinned to the reference's compiled cpu_toed.cpp by the golden tests).
Tolerances are stated in tests/parity.py."""
import glob
import os
import sys

import numpy as np
import pytest

from parity import assert_parity, compare

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

GOLDEN = sorted(glob.glob(os.path.join(os.path.dirname(__file__), "golden", "toed_ref_*.npz")))


def _run(detector, img, sigma=2.0, dtype=None, **kw):
    t = torch.from_numpy(np.ascontiguousarray(img)).cuda()
    if dtype is not None:
        t = t.to(dtype)
    e, i = detector.detect(t, sigma=sigma, with_index=True, **kw)
    torch.cuda.synchronize()
    return e.cpu().numpy(), i.cpu().numpy()


def _report(name, rep):
    print("%s: %s" % (name, rep))
    sys.stdout.flush()


@pytest.mark.parametrize("path", GOLDEN, ids=[os.path.basename(p)[9:-4] for p in GOLDEN])
def test_golden_inputs(detector, oracle, path):
    d = np.load(path)
    img = d["img"]
    ref = oracle.detect(img.astype(np.float64), 2.0, margins=True)
    assert ref["n"] == len(d["edges"])
    e, i = _run(detector, img)
    rep = compare(ref, e, i)
    _report(os.path.basename(path), rep)
    assert_parity(rep)


@pytest.mark.parametrize("sigma", [1.0, 1.5, 2.0])
@pytest.mark.parametrize("shape", [(480, 752), (301, 413), (64, 40)])
def test_seeded_scenes_sigma(detector, oracle, scene, sigma, shape):
    H, W = shape
    img = scene.shapes_scene(100 + H, H, W)
    ref = oracle.detect(img.astype(np.float64), sigma, margins=True)
    e, i = _run(detector, img, sigma)
    rep = compare(ref, e, i)
    _report("scene %dx%d sigma %g" % (H, W, sigma), rep)
    assert_parity(rep)


@pytest.mark.parametrize("sigma", [1.0, 1.5, 2.0])
def test_orientation_at_subpixel_location(pkg, oracle, scene, sigma):
    """Orientation mode 1 (the MATLAB detector's QaD path) against the
    oracle's detect_qad: the edgels, positions and strengths of mode 0, the
    orientation from the responses interpolated at the subpixel location
    within the parity tolerances; the batch plan gives the same."""
    img = scene.shapes_scene(700 + int(10 * sigma), 160, 224)
    det = pkg.Detector(0)
    det.set_orientation_mode(1)
    e1, i1 = _run(det, img, sigma)
    b = pkg.BatchDetector(det, [torch.from_numpy(img).cuda()], [sigma])
    b.run(graph=True)
    eb = b.results()[0].cpu().numpy()
    b.close()
    det.set_orientation_mode(0)
    e0, i0 = _run(det, img, sigma)
    det.close()
    assert np.array_equal(eb, e1)
    assert np.array_equal(i1, i0) and np.array_equal(e1[:, [0, 1, 3]], e0[:, [0, 1, 3]])
    ref = oracle.detect_qad(img.astype(np.float64), sigma, margins=True)
    rep = compare(ref, e1, i1)
    _report("QaD orientation, scene 160x224 sigma %g" % sigma, rep)
    assert_parity(rep)
    moved = np.abs(np.angle(np.exp(1j * (e1[:, 2] - e0[:, 2]))))
    print("orientation change from mode 0: max %.3g rad, median %.3g rad" % (moved.max(), np.median(moved)))
    assert moved.max() > 1e-4


def test_full_frame_2160x3840(detector, oracle, scene):
    # the plain workload's own frame (96 shapes, 2160 x 3840 u8, sigma 2), edgel for edgel
    img = scene.shapes_scene(2160 * 3840, 2160, 3840)
    ref = oracle.detect(img.astype(np.float64), 2.0, margins=True)
    e, i = _run(detector, img)
    rep = compare(ref, e, i)
    _report("scene 2160x3840 sigma 2", rep)
    assert ref["n"] > 100000
    assert_parity(rep)


def test_float_input_matches_u8(detector, oracle, scene):
    img = scene.shapes_scene(21, 200, 260)
    e8, i8 = _run(detector, img)
    ef, i_f = _run(detector, img, dtype=torch.float32)
    assert np.array_equal(i8, i_f)
    assert np.array_equal(e8, ef)


def test_orientation_at_subpixel_location_float_input(pkg, scene):
    """Orientation mode 1 reads float images through its own load path; on
    integer-valued pixels it gives the u8 result bit for bit (NMS border at its
    minimum of one pixel, so that windows near the image edge are clipped)."""
    img = scene.shapes_scene(23, 150, 190)
    det = pkg.Detector(0)
    det.set_orientation_mode(1)
    try:
        e8, i8 = _run(det, img, 1.5, border=1)
        ef, i_f = _run(det, img, 1.5, dtype=torch.float32, border=1)
    finally:
        det.close()
    assert len(e8) > 100
    assert np.array_equal(i8, i_f)
    assert np.array_equal(e8, ef)


def test_float_input_non_integer(detector, oracle, scene):
    img = scene.shapes_scene(22, 160, 200).astype(np.float32) * np.float32(0.731) + np.float32(3.3)
    ref = oracle.detect(img.astype(np.float64), 1.5, margins=True)
    e, i = _run(detector, img, 1.5)
    rep = compare(ref, e, i)
    _report("float scene", rep)
    assert_parity(rep)


def test_strided_rows(detector, oracle, scene):
    img = scene.shapes_scene(23, 128, 190)
    big = torch.zeros((128, 256), dtype=torch.uint8, device="cuda")
    big[:, :190] = torch.from_numpy(img).cuda()
    view = big[:, :190]
    e, i = detector.detect(view, with_index=True)
    e2, i2 = _run(detector, img)
    assert np.array_equal(i.cpu().numpy(), i2) and np.array_equal(e.cpu().numpy(), e2)


@pytest.mark.parametrize("shape", [(1, 1), (5, 7), (10, 300), (11, 11), (33, 17)])
def test_tiny_and_ragged(detector, oracle, shape):
    rng = np.random.default_rng(shape[0] * 1000 + shape[1])
    img = rng.integers(0, 256, shape, dtype=np.uint8)
    ref = oracle.detect(img.astype(np.float64), 2.0, margins=True)
    e, i = _run(detector, img)
    rep = compare(ref, e, i)
    _report("tiny %s" % (shape,), rep)
    assert_parity(rep)
    if shape[0] <= 10 or shape[1] <= 10:
        assert len(i) == 0


def test_constant_and_step(detector):
    e, i = _run(detector, np.full((64, 96), 128, np.uint8))
    assert len(i) == 0
    img = np.zeros((64, 96), np.uint8)
    img[:, 48:] = 200
    e, i = _run(detector, img)
    IW = 192
    cols = i % IW
    assert len(i) > 0 and np.all(np.abs(e[:, 0] - 47.0) < 0.05) and np.all(np.abs(cols - 95) <= 1)


def test_capacity_overflow_reports_exact_count(detector, pkg, scene):
    img = torch.from_numpy(scene.shapes_scene(5, 200, 200)).cuda()
    full = detector.detect(img)
    n = full.shape[0]
    assert n > 10
    with pytest.raises(pkg.ToedError) as ei:
        detector.detect(img, capacity=n // 2)
    assert ei.value.code == 5
    # the rows that fit are the first rows of the full list
    edges = torch.zeros((n // 2, 4), dtype=torch.float32, device="cuda")
    counts = torch.zeros(2, dtype=torch.int32, device="cuda")
    detector.detect_into(img, edges, counts)
    c = counts.cpu().numpy()
    assert c[0] == n and c[1] & 1
    assert torch.equal(edges, full[: n // 2])


def test_staging_overflow_is_flagged(pkg, scene):
    det = pkg.Detector(0, staging_fraction=0.002)
    img = torch.from_numpy(scene.shapes_scene(6, 200, 200)).cuda()
    with pytest.raises(pkg.ToedError) as ei:
        det.detect(img)
    assert ei.value.code == 6
    det.close()


def test_tile_rows_variants_identical(pkg, scene):
    img = torch.from_numpy(scene.shapes_scene(8, 250, 333)).cuda()
    outs = []
    for tr in (8, 16):
        det = pkg.Detector(0, tile_rows=tr)
        e, i = det.detect(img, with_index=True)
        outs.append((e.cpu().numpy(), i.cpu().numpy()))
        det.close()
    assert np.array_equal(outs[0][1], outs[1][1])
    assert np.array_equal(outs[0][0], outs[1][0])


def test_bad_arguments(detector, pkg):
    img = torch.zeros((32, 32), dtype=torch.uint8, device="cuda")
    with pytest.raises(pkg.ToedError):
        detector.detect(img, sigma=2.7)
    with pytest.raises(TypeError):
        detector.detect(img.to(torch.int16))
    with pytest.raises(pkg.ToedError):
        detector.detect(img, border=0)


def test_reference_class_mirror_double(pkg, scene, tmp_path):
    """The T = double instantiation (main_gpu_cpu.cpp:200-209): the same
    detection, its FP32 rows widened into the double buffer."""
    img = scene.shapes_scene(31, 120, 150)
    out = {}
    for dt in (np.float32, np.float64):
        cls = pkg.ThirdOrderEdgeDetectionGPU(0, 120, 150, 17, 2, dtype=dt)
        cls.set_output_dir(str(tmp_path))
        cls.preprocessing(img)
        cls.convolve_img()
        buf = np.zeros((4 * 120 * 150, 4), dt)
        n = cls.non_maximum_suppresion(buf)
        assert cls.subpix_edge_pts_final.dtype == dt
        out[dt] = buf[:n]
    assert len(out[np.float32]) > 100
    assert np.array_equal(out[np.float64], out[np.float32].astype(np.float64))
    with pytest.raises(NotImplementedError):
        pkg.ThirdOrderEdgeDetectionGPU(0, 120, 150, 17, 2, dtype=np.float16)


def test_reference_class_mirror(pkg, oracle, scene, tmp_path):
    img = scene.shapes_scene(31, 120, 150)
    cls = pkg.ThirdOrderEdgeDetectionGPU(0, 120, 150, 17, 2)
    cls.set_output_dir(str(tmp_path))
    cls.preprocessing(img)
    cls.convolve_img()
    TOED_edges = np.zeros((4 * 120 * 150, 4), np.float32)
    n = cls.non_maximum_suppresion(TOED_edges)
    assert n == cls.edge_pt_list_idx > 0
    assert cls.time_conv > 0 and cls.time_nms > 0
    txt = np.loadtxt(tmp_path / "data_final_output_gpu.txt")
    np.testing.assert_allclose(txt, TOED_edges[:n], rtol=1e-6, atol=1e-6)
    # the byte-stream preprocessing overload gives the same result
    import io
    cls2 = pkg.ThirdOrderEdgeDetectionGPU(0, 120, 150, 17, 2)
    cls2.set_output_dir(str(tmp_path))
    cls2.preprocessing(io.BytesIO(img.tobytes()))
    cls2.convolve_img()
    e2 = np.zeros_like(TOED_edges)
    assert cls2.non_maximum_suppresion(e2) == n
    assert np.array_equal(e2[:n], TOED_edges[:n])
    with pytest.raises(ValueError):
        pkg.ThirdOrderEdgeDetectionGPU(0, 10, 10, 19, 2)


def test_unfused_launch_layer_f64(pkg, oracle, scene):
    """toed_gpu_convolve_sigma_f64 / toed_gpu_nms_thresh_f64 (the T = double
    form of the launch layer, taps from sigma) against the oracle's FP64 planes
    and list:
    planes to 1e-9 of their range, the same edgel set apart from decisions
    within 1e-9 of a tie, positions to 1e-8 interp px."""
    import ctypes
    L = pkg._lib.lib()
    H, W = 90, 120
    img = scene.shapes_scene(41, H, W)
    t = torch.from_numpy(img.astype(np.float64)).cuda()
    planes = [torch.zeros((2 * H, 2 * W), dtype=torch.float64, device="cuda") for _ in range(4)]
    ptr = lambda x: ctypes.c_void_p(x.data_ptr())
    assert L.toed_gpu_convolve_sigma_f64(0, H, W, 2 * H, 2 * W, ptr(t), *[ptr(p) for p in planes], 2.0) == 0
    torch.cuda.synchronize()
    c = oracle.convolve(img.astype(np.float64), 2.0)
    Ix, Iy, mag, ori = (p.cpu().numpy() for p in planes)
    for got, key in ((Ix, "Ix"), (Iy, "Iy"), (mag, "mag")):
        assert np.abs(got - c[key]).max() < 1e-9 * np.abs(c[key]).max(), key
    strong = c["mag"] > 2
    d = np.abs(ori - c["orient"])[strong]
    d = np.minimum(d, np.abs(d - 2 * np.pi))
    assert np.nanmax(d) < 1e-6
    px = torch.zeros((2 * H, 2 * W), dtype=torch.float64, device="cuda")
    py = torch.zeros_like(px)
    assert L.toed_gpu_nms_thresh_f64(0, 2 * H, 2 * W, ptr(planes[0]), ptr(planes[1]), ptr(planes[2]), ptr(px),
                                     ptr(py), 2.0, 10) == 0
    torch.cuda.synchronize()
    pxm, pym = px.cpu().numpy().reshape(-1), py.cpu().numpy().reshape(-1)
    sel = np.flatnonzero(pxm)
    ref = oracle.detect(img.astype(np.float64), 2.0, margins=True)
    only = np.setxor1d(sel, ref["eidx"])
    mm, ms = ref["mmag"].reshape(-1), ref["msub"].reshape(-1)
    assert np.all((mm[only] < 1e-9) | (ms[only] < 1e-9))
    assert len(sel) > 100 and len(only) <= 2
    common, ia, _ = np.intersect1d(ref["eidx"], sel, return_indices=True)
    e = ref["edges"][ia]
    assert np.abs(pxm[common] - (2 * e[:, 0] + 1)).max() < 1e-8
    assert np.abs(pym[common] - (2 * e[:, 1] + 1)).max() < 1e-8


def test_unfused_launch_layer(pkg, oracle, scene):
    """toed_gpu_convolve_sigma_f32 / toed_gpu_nms_thresh_f32 (the launch layer
    with taps from sigma) against the oracle's planes and its subpixel maps."""
    import ctypes
    L = pkg._lib.lib()
    H, W = 90, 120
    img = scene.shapes_scene(41, H, W)
    t = torch.from_numpy(img.astype(np.float32)).cuda()
    planes = [torch.zeros((2 * H, 2 * W), dtype=torch.float32, device="cuda") for _ in range(4)]
    ptr = lambda x: ctypes.c_void_p(x.data_ptr())
    assert L.toed_gpu_convolve_sigma_f32(0, H, W, 2 * H, 2 * W, ptr(t), *[ptr(p) for p in planes], 2.0) == 0
    torch.cuda.synchronize()
    c = oracle.convolve(img.astype(np.float64), 2.0)
    Ix, Iy, mag, ori = (p.cpu().numpy().astype(np.float64) for p in planes)
    assert np.abs(Ix - c["Ix"]).max() < 1e-3 and np.abs(Iy - c["Iy"]).max() < 1e-3
    assert np.abs(mag - c["mag"]).max() < 1e-3
    strong = c["mag"] > 2
    d = np.abs(ori - c["orient"])[strong]
    d = np.minimum(d, np.abs(d - 2 * np.pi))
    assert np.nanquantile(d, 0.99) < 1e-2
    px = torch.zeros((2 * H, 2 * W), dtype=torch.float32, device="cuda")
    py = torch.zeros_like(px)
    assert L.toed_gpu_nms_thresh_f32(0, 2 * H, 2 * W, ptr(planes[0]), ptr(planes[1]), ptr(planes[2]), ptr(px),
                                     ptr(py), 2.0, 10) == 0
    torch.cuda.synchronize()
    sel = np.flatnonzero(px.cpu().numpy().reshape(-1))
    ref = oracle.detect(img.astype(np.float64), 2.0, margins=True)
    only = np.setxor1d(sel, ref["eidx"])
    mm, ms = ref["mmag"].reshape(-1), ref["msub"].reshape(-1)
    assert np.all((mm[only] < 2e-3) | (ms[only] < 2e-3))


@pytest.mark.parametrize("dtype", ["float32", "float64"])
def test_launch_layer_reference_arguments(pkg, oracle, scene, dtype):
    """toed_gpu_convolve_f32/_f64 and toed_gpu_nms_f32/_f64 called with the
    reference's own argument lists (gpu_kernels.hpp:24-63), the way
    gpu_toed.hpp:379-394 and :420 call them: eight device tap arrays (Gx, Gxx,
    Gxxx, G_of_x, then the shifted four; the oracle's sigma = 2 taps, which
    test_oracle pins to the reference's literals) and an NMS with the
    reference's threshold 2 and border 10 built in.  Planes match the oracle's
    (FP32: to 1e-3; FP64: to 1e-9 of their range) and the maps the sigma /
    threshold forms' bit for bit."""
    import ctypes
    L = pkg._lib.lib()
    H, W = 90, 120
    img = scene.shapes_scene(43, H, W)
    tdt = getattr(torch, dtype)
    t = torch.from_numpy(img.astype(dtype)).cuda()
    _, _, taps = oracle.taps(2.0)
    order = [1, 2, 3, 0, 5, 6, 7, 4]       # Gx, Gxx, Gxxx, G_of_x, Gx_sh, Gxx_sh, Gxxx_sh, G_of_x_sh
    dev_taps = [torch.from_numpy(np.ascontiguousarray(taps[k], dtype)).cuda() for k in order]
    assert all(d.numel() == 19 for d in dev_taps)
    ptr = lambda x: ctypes.c_void_p(x.data_ptr())
    planes = [torch.zeros((2 * H, 2 * W), dtype=tdt, device="cuda") for _ in range(4)]
    conv = L.toed_gpu_convolve_f32 if dtype == "float32" else L.toed_gpu_convolve_f64
    nms = L.toed_gpu_nms_f32 if dtype == "float32" else L.toed_gpu_nms_f64
    assert conv(0, H, W, 2 * H, 2 * W, ptr(t), *[ptr(p) for p in planes], *[ptr(d) for d in dev_taps]) == 0
    px = torch.zeros((2 * H, 2 * W), dtype=tdt, device="cuda")
    py = torch.zeros_like(px)
    assert nms(0, 2 * H, 2 * W, ptr(planes[0]), ptr(planes[1]), ptr(planes[2]), ptr(px), ptr(py)) == 0
    torch.cuda.synchronize()
    c = oracle.convolve(img.astype(np.float64), 2.0)
    tol = 1e-3 if dtype == "float32" else 1e-9
    for k, key in enumerate(("Ix", "Iy", "mag")):
        got = planes[k].cpu().numpy().astype(np.float64)
        scale = 1.0 if dtype == "float32" else np.abs(c[key]).max()
        assert np.abs(got - c[key]).max() < tol * scale, key
    # the sigma / threshold forms on the same image: the same maps
    sp = [torch.zeros_like(p) for p in planes]
    sconv = L.toed_gpu_convolve_sigma_f32 if dtype == "float32" else L.toed_gpu_convolve_sigma_f64
    snms = L.toed_gpu_nms_thresh_f32 if dtype == "float32" else L.toed_gpu_nms_thresh_f64
    assert sconv(0, H, W, 2 * H, 2 * W, ptr(t), *[ptr(p) for p in sp], 2.0) == 0
    qx, qy = torch.zeros_like(px), torch.zeros_like(py)
    assert snms(0, 2 * H, 2 * W, ptr(sp[0]), ptr(sp[1]), ptr(sp[2]), ptr(qx), ptr(qy), 2.0, 10) == 0
    torch.cuda.synchronize()
    sel = np.flatnonzero(px.cpu().numpy().reshape(-1))
    assert len(sel) > 100
    assert np.array_equal(sel, np.flatnonzero(qx.cpu().numpy().reshape(-1)))
    np.testing.assert_allclose(px.cpu().numpy(), qx.cpu().numpy(), rtol=0, atol=1e-9 if dtype == "float64" else 1e-4)
    # the edgel set is the oracle's (decisions within the FP32 margin aside)
    ref = oracle.detect(img.astype(np.float64), 2.0, margins=True)
    only = np.setxor1d(sel, ref["eidx"])
    mm, ms = ref["mmag"].reshape(-1), ref["msub"].reshape(-1)
    assert np.all((mm[only] < 2e-3) | (ms[only] < 2e-3))
