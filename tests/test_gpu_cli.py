"""The command-line drop-in (toed_detect_cli) on the GPU: the reference's argv
and output file (main_gpu_cpu.cpp:103-153, gpu_toed.hpp:452-478), and the
batch mode that replaces run_gpu_batch.sh's process per image; and the
header-only C++ class driven as main_gpu_cpu.cpp drives the reference's; all
against the library's own single-image detection."""
import os
import subprocess

import numpy as np
import pytest

from test_image_io import _pgm, _png

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

CLI = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "third-order-edge-detector_amd",
                   "toed_detect_cli")
DRIVER = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "toed_class_driver")


def _read_list(path):
    rows = [l.rstrip("\n").split("\t") for l in open(path) if l.strip()]
    for r in rows:
        assert r[-1] == "" and len(r) == 5        # a tab after every value
    return np.array([[float(v) for v in r[:4]] for r in rows], np.float64).reshape(-1, 4)


def _run(args, timeout=120):
    print("cli:", " ".join(args), flush=True)
    return subprocess.run([CLI] + args, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                          timeout=timeout)


def test_single_image_reference_argv(pkg, detector, scene, tmp_path):
    assert os.path.exists(CLI), "toed_detect_cli is not built"
    img = scene.shapes_scene(321, 240, 320)
    p = str(tmp_path / "scene.pgm")
    _pgm(p, img)
    out_dir = str(tmp_path / "out" / "nested")
    r = _run([p, "4", "0", out_dir])
    print(r.stdout)
    assert r.returncode == 0
    got = _read_list(os.path.join(out_dir, "data_final_output_gpu.txt"))
    want = detector.detect(torch.from_numpy(img).cuda()).cpu().numpy().astype(np.float64)
    assert "Number of GPU edges = %d" % len(want) in r.stdout
    assert got.shape == want.shape and len(want) > 100
    # precision 10 text of FP32 values
    np.testing.assert_allclose(got, want, rtol=1e-9, atol=1e-9)


def test_missing_file_and_usage(tmp_path):
    r = _run([str(tmp_path / "nope.pgm")])
    assert r.returncode == 0 and "not found" in r.stdout     # main_gpu_cpu.cpp:114-120
    r = _run([])
    assert r.returncode == 0 and "Usage" in r.stdout


def test_batch_mode_matches_single_images(pkg, detector, scene, tmp_path):
    imgs = {"a_scene": scene.shapes_scene(401, 200, 300), "b_scene": scene.shapes_scene(402, 480, 640),
            "c_scene": scene.shapes_scene(403, 150, 100)}
    paths = []
    for k, (name, im) in enumerate(imgs.items()):
        p = str(tmp_path / (name + (".png" if k == 1 else ".pgm")))
        _png(p, im, 0) if k == 1 else _pgm(p, im)
        paths.append(p)
    out_dir = str(tmp_path / "batch_out")
    r = _run(["--batch", "--out", out_dir, "--autotune"] + paths)
    print(r.stdout)
    assert r.returncode == 0
    for name, im in imgs.items():
        got = _read_list(os.path.join(out_dir, name + "_edges_gpu.txt"))
        want = detector.detect(torch.from_numpy(im).cuda()).cpu().numpy().astype(np.float64)
        assert got.shape == want.shape and len(want) > 10, name
        np.testing.assert_allclose(got, want, rtol=1e-9, atol=1e-9)


def test_reference_pgm_byte_shift(pkg, detector, scene, tmp_path):
    """TOED_PGM_RAW_OFFSET=1 reads the PGM as the reference's no-OpenCV build
    does (the header's last whitespace byte becomes pixel 0, every pixel one
    position late), and the edgels are those of that shifted image."""
    img = scene.shapes_scene(404, 120, 160)
    p = str(tmp_path / "shift.pgm")
    _pgm(p, img, header=b"P5\n%d %d\n255\n")
    out_dir = str(tmp_path / "out_shift")
    env = dict(os.environ, TOED_PGM_RAW_OFFSET="1")
    r = subprocess.run([CLI, p, "1", "0", out_dir], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                       timeout=120, env=env)
    print(r.stdout)
    assert r.returncode == 0
    shifted = np.concatenate([[ord("\n")], img.reshape(-1)[:-1]]).astype(np.uint8).reshape(img.shape)
    want = detector.detect(torch.from_numpy(shifted).cuda()).cpu().numpy().astype(np.float64)
    got = _read_list(os.path.join(out_dir, "data_final_output_gpu.txt"))
    assert got.shape == want.shape
    np.testing.assert_allclose(got, want, rtol=1e-9, atol=1e-9)


def test_header_class_float_and_double(pkg, detector, scene, tmp_path):
    """ThirdOrderEdgeDetectionGPU<T> from include/toed_gpu.hpp, run as
    main_gpu_cpu.cpp:195-209 runs the reference's class: T = float returns the
    library's own list, T = double the same rows widened, and each writes its
    text list."""
    assert os.path.exists(DRIVER), "tests/cpp/toed_class_driver is not built"
    img = scene.shapes_scene(405, 180, 240)
    raw = tmp_path / "img.raw"
    raw.write_bytes(img.tobytes())
    for t in ("f32", "f64"):
        (tmp_path / t).mkdir()
    print("driver:", DRIVER, flush=True)
    r = subprocess.run([DRIVER, str(raw), "180", "240", "2", str(tmp_path)], stdout=subprocess.PIPE,
                       stderr=subprocess.STDOUT, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0
    want = detector.detect(torch.from_numpy(img).cuda()).cpu().numpy()
    assert len(want) > 100
    assert "f32 edges = %d" % len(want) in r.stdout and "f64 edges = %d" % len(want) in r.stdout
    f32 = np.fromfile(str(tmp_path / "edges_f32.bin"), np.float32).reshape(-1, 4)
    f64 = np.fromfile(str(tmp_path / "edges_f64.bin"), np.float64).reshape(-1, 4)
    assert np.array_equal(f32, want)
    assert np.array_equal(f64, want.astype(np.float64))
    for t in ("f32", "f64"):
        got = _read_list(str(tmp_path / t / "data_final_output_gpu.txt"))
        np.testing.assert_allclose(got, want.astype(np.float64), rtol=1e-9, atol=1e-9)


GOLDEN_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
GOLDEN_DET = sorted(__import__("glob").glob(os.path.join(GOLDEN_DIR, "toed_ref_*.npz")))
GOLDEN_SP = sorted(__import__("glob").glob(os.path.join(GOLDEN_DIR, "subpix_ref_*.npz")))
SP_DRIVER = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "toed_subpix_driver")


@pytest.mark.parametrize("path", GOLDEN_DET, ids=[os.path.basename(p)[9:-4] for p in GOLDEN_DET])
def test_cli_against_reference_golden_lists(oracle, path, tmp_path):
    """toed_detect_cli with the reference binary's argv on each image of the
    golden set: its data_final_output_gpu.txt against the list the reference's
    own cpu_toed.cpp wrote for that image (tests/golden).  Rows are matched by
    position (within 1e-2 px) and must keep the reference's order; matched rows
    meet the parity tolerances (1e-3 px, 1e-3 rad, 1e-4 relative strength); a
    row found on one side only must sit at a pixel whose threshold or peak-test
    margin in the FP64 oracle is below 2e-3 (the FP32 filters can decide such a
    pixel either way)."""
    from scipy.spatial import cKDTree
    d = np.load(path)
    img, gold = d["img"], d["edges"]
    H, W = img.shape
    p = str(tmp_path / "img.pgm")
    _pgm(p, img)
    out_dir = str(tmp_path / "out")
    r = _run([p, "1", "0", out_dir])
    assert r.returncode == 0, r.stdout
    got = _read_list(os.path.join(out_dir, "data_final_output_gpu.txt"))
    assert "Number of GPU edges = %d" % len(got) in r.stdout
    ref = oracle.detect(img.astype(np.float64), 2.0, margins=True)
    assert ref["n"] == len(gold)
    np.testing.assert_allclose(ref["edges"][:, [0, 1, 3]], gold[:, [0, 1, 3]], rtol=1e-9, atol=1e-9)
    dist, j = cKDTree(gold[:, :2]).query(got[:, :2], distance_upper_bound=1e-2)
    ok = np.isfinite(dist)
    jm = j[ok]
    assert np.all(np.diff(jm) > 0), "CLI rows out of the reference's order"
    a, b = gold[jm], got[ok]
    dxy = np.abs(a[:, :2] - b[:, :2]).max()
    both = ~np.isnan(a[:, 2]) & ~np.isnan(b[:, 2])
    assert np.array_equal(np.isnan(a[:, 2]), np.isnan(b[:, 2]))
    dth = np.abs(np.angle(np.exp(1j * (a[both, 2] - b[both, 2])))).max() if both.any() else 0.0
    ds = (np.abs(a[:, 3] - b[:, 3]) / np.abs(a[:, 3])).max()
    mm, ms = ref["mmag"].reshape(-1), ref["msub"].reshape(-1)
    IW = 2 * W
    only_gold = np.setdiff1d(np.arange(len(gold)), jm)
    gold_expl = (mm[ref["eidx"][only_gold]] < 2e-3) | (ms[ref["eidx"][only_gold]] < 2e-3)
    # a CLI-only row: the interp pixels around its subpixel position
    only_cli = got[~ok]
    cli_expl = []
    for x, y in only_cli[:, :2]:
        C, R = int(round(2 * x + 1)), int(round(2 * y + 1))
        win = [(R + dr) * IW + (C + dc) for dr in (-2, -1, 0, 1, 2) for dc in (-2, -1, 0, 1, 2)
               if 0 <= R + dr < 2 * H and 0 <= C + dc < IW]
        cli_expl.append(bool(((mm[win] < 2e-3) | (ms[win] < 2e-3)).any()))
    print("%s: %d reference rows, %d CLI rows, %d matched, only reference %d, only CLI %d, max |dxy| %.2e px, "
          "|dtheta| %.2e rad, strength %.2e rel" % (os.path.basename(path), len(gold), len(got), ok.sum(),
                                                   len(only_gold), len(only_cli), dxy, dth, ds), flush=True)
    assert ok.sum() >= 0.99 * len(gold) and len(gold) > 100
    assert gold_expl.all() and all(cli_expl)
    assert dxy <= 1e-3 and dth <= 1e-3 and ds <= 1e-4


@pytest.mark.parametrize("T", ["double", "float"])
@pytest.mark.parametrize("path", GOLDEN_SP, ids=[os.path.basename(p)[11:-4] for p in GOLDEN_SP])
def test_cpp_subpix_class_against_reference_golden(path, T, tmp_path):
    """The header-only C++ SubpixelTOCorrectionGPU<T> (T = double as the
    reference's driver instantiates it, and float), driven as
    main_subpix_to_correction.cpp:279-293 drives SubpixelTOCorrectionCPU<double>,
    on each golden map (constructed with the fixture's n = interp_n): the
    edgel list equals the reference's own list (same edgels, in order) within
    1e-3 px, 1e-3 rad and 1e-4 relative strength;
    TO_edgemap / TO_orientation have the reference's support and values within
    the same tolerances; the text list is written."""
    assert os.path.exists(SP_DRIVER), "toed_subpix_driver is not built"
    z = np.load(path)
    thr, sig, nbr = (float(v) for v in z["params"])
    E = np.ascontiguousarray(z["E"], np.float64)
    H, W = E.shape
    mp = str(tmp_path / "map.f64")
    E.tofile(mp)
    op = "-"
    if "theta" in z.files:
        op = str(tmp_path / "orient.f64")
        np.ascontiguousarray(z["theta"], np.float64).tofile(op)
    out, od = str(tmp_path / "out.bin"), str(tmp_path / "o")
    os.makedirs(od)
    interp_n = int(z["interp_n"]) if "interp_n" in z.files else 0
    args = [SP_DRIVER, mp, op, str(H), str(W), repr(thr), repr(sig), repr(nbr), T, out, od, str(interp_n)]
    print("driver:", " ".join(args[3:9] + args[11:]), flush=True)
    r = subprocess.run(args, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=120)
    assert r.returncode == 0, r.stdout
    with open(out, "rb") as f:
        n = int(np.frombuffer(f.read(8), np.int64)[0])
        e = np.frombuffer(f.read(n * 32), np.float64).reshape(n, 4)
        te = np.frombuffer(f.read(H * W * 8), np.float64).reshape(H, W)
        to = np.frombuffer(f.read(H * W * 8), np.float64).reshape(H, W)
    gold = z["edges"]
    assert "Number of corrected edgels = %d" % n in r.stdout
    assert n == len(gold) > 100
    assert os.path.exists(os.path.join(od, "TO_corrected_edges.txt"))
    dxy = np.abs(e[:, :2] - gold[:, :2]).max()
    dth = np.abs(np.angle(np.exp(1j * (e[:, 2] - gold[:, 2])))).max()
    ds = (np.abs(e[:, 3] - gold[:, 3]) / np.abs(gold[:, 3])).max()
    print("%s <%s>: %d edgels, max |dxy| %.2e px, |dtheta| %.2e rad, strength %.2e rel"
          % (os.path.basename(path), T, n, dxy, dth, ds), flush=True)
    assert dxy <= 1e-3 and dth <= 1e-3 and ds <= 1e-4
    assert np.array_equal(te != 0, z["TO_E"] != 0)
    sup = z["TO_E"] != 0
    assert (np.abs(te[sup] - z["TO_E"][sup]) / np.abs(z["TO_E"][sup])).max() <= 1e-4
    assert np.abs(np.angle(np.exp(1j * (to[sup] - z["TO_O"][sup])))).max() <= 1e-3
