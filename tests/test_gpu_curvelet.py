"""Curvelet formation downstream of the GPU detector, as the reference's
driver runs it on its FP64 GPU list (main_gpu_cpu.cpp:211-281): the list the
HIP path produces, widened to double, grouped by the library and by the
reference's own curvelet code (compiled into oracle/_ref) must give the same
outputs bit for bit; the CLI's --curvelets files must be the reference
writer's text of those outputs (main_gpu_cpu.cpp:69-97)."""
import os
import subprocess

import numpy as np
import pytest

from test_image_io import _pgm

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

CLI = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "third-order-edge-detector_amd",
                   "toed_detect_cli")


def _ref_text(a):
    # std::ostream defaults (precision 6, %g style), a tab after every value
    fmt = "%d" if a.dtype.kind == "i" else "%g"
    return "".join("".join((fmt % v) + "\t" for v in row) + "\n" for row in a)


def _same(a, b):
    assert a.shape == b.shape
    if a.dtype.kind == "f":
        np.testing.assert_array_equal(a.view(np.int64), b.view(np.int64))
    else:
        np.testing.assert_array_equal(a, b)


def test_gpu_list_curvelets_match_reference(pkg, oracle, detector, scene):
    if not oracle.curvelet_ref_available():
        pytest.skip("the compiled reference curvelet code is not present")
    H, W = 480, 752
    img = scene.shapes_scene(3, H, W)
    print("case: shapes_scene(3, 480, 752) -> detect -> curvelets", flush=True)
    rows = detector.detect(torch.from_numpy(img).cuda()).cpu().numpy().astype(np.float64)
    assert len(rows) > 1000
    _, ref = oracle.run_curvelet_ref(rows, H, W)
    for t in (0, 1, 2):
        chain, info = pkg.form_curvelets(rows, H, W, output_type=t)
        _same(chain, ref[t][0])
        _same(info, ref[t][1])


def test_cli_curvelets_files(pkg, oracle, scene, tmp_path):
    assert os.path.exists(CLI), "toed_detect_cli is not built"
    H, W = 400, 400
    img = scene.shapes_scene(17, H, W)
    p = str(tmp_path / "scene.pgm")
    _pgm(p, img)
    out_dir = str(tmp_path / "out")
    print("cli: --curvelets on shapes_scene(17, 400, 400)", flush=True)
    r = subprocess.run([CLI, p, "1", "0", out_dir, "--curvelets"], stdout=subprocess.PIPE,
                       stderr=subprocess.STDOUT, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0
    # the list file holds the FP32 rows at precision 10, which parses back exactly
    txt = [l.rstrip("\n").split("\t")[:4] for l in open(os.path.join(out_dir, "data_final_output_gpu.txt"))]
    rows = np.array(txt, dtype=np.float32).astype(np.float64).reshape(-1, 4)
    chain, info = pkg.form_curvelets(rows, H, W)
    assert "Curvelet formation: %d curvelets" % len(chain) in r.stdout
    assert open(os.path.join(out_dir, "chain.txt")).read() == _ref_text(chain)
    assert open(os.path.join(out_dir, "info.txt")).read() == _ref_text(info)
    if oracle.curvelet_ref_available():
        _, ref = oracle.run_curvelet_ref(rows, H, W, types="0")
        _same(chain, ref[0][0])
        _same(info, ref[0][1])
