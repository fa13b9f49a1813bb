"""Curvelet formation (host code in libtoed_hip.so) against the reference's
own curvelet code (curvelet/*.hpp, compiled by oracle/Makefile): bit for bit,
through the golden hashes it wrote (tests/golden/make_curvelet_golden.py)
and, where the compiled reference is present, directly.  CPU only."""
import hashlib
import json
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")
sys.path.insert(0, GOLD)
from make_curvelet_golden import cases  # noqa: E402

TABLE = json.load(open(os.path.join(GOLD, "curvelet_ref.json")))
CASES = {name: (e, H, W, v) for name, e, H, W, v in cases()}


def digest(a):
    return hashlib.sha256(np.asfortranarray(a).tobytes(order="F")).hexdigest()


def run(pkg, e, H, W, p, nthreads=0):
    centered, bidirectional = pkg.curvelet.style_flags(p["style"])
    cf = pkg.CurveletFormation(e, H, W, p["rad"], p["gap"], p["dx"], p["dt"], p["token_len"], p["max_k"],
                               p["group"], centered, bidirectional)
    cf.execute(nthreads)
    return {t: cf.get_output_arrary(t) for t in (0, 1, 2)}


@pytest.mark.parametrize("name", sorted(CASES))
def test_curvelets_match_reference_golden(pkg, name):
    e, H, W, _ = CASES[name]
    rec = TABLE[name]
    out = run(pkg, e, H, W, rec["params"])
    for t in (0, 1, 2):
        chain, info = out[t]
        want = rec[str(t)]
        assert [chain.shape[0], chain.shape[1], info.shape[1]] == want["shape"], (name, t)
        assert chain.dtype == np.int32 and info.dtype == np.float64
        assert digest(chain) == want["chain_sha256"], (name, t, "chain")
        assert digest(info) == want["info_sha256"], (name, t, "info")


def test_full_outputs_smallest_case(pkg):
    d = np.load(os.path.join(GOLD, "curvelet_ref_scene_64x96_s1.npz"))
    e, H, W, _ = CASES["scene_64x96_s1"]
    chain, info = pkg.form_curvelets(e, H, W)            # the drivers' settings, output type 0
    np.testing.assert_array_equal(chain, d["chain0"])
    np.testing.assert_array_equal(info.view(np.int64), d["info0"].view(np.int64))
    # column 0 is the 1-based anchor id, column 1 the anchor itself (anchor-leading style)
    assert (chain[:, 0] == chain[:, 1]).all()
    assert (np.diff(chain[:, 0]) >= 0).all()                # anchors in id order
    for t in (1, 2):
        c, i = pkg.form_curvelets(e, H, W, output_type=t)
        np.testing.assert_array_equal(c, d["chain%d" % t])
        np.testing.assert_array_equal(i.view(np.int64), d["info%d" % t].view(np.int64))


def test_reference_build_stops_on_polyarc_output():
    # recorded by the golden script: the reference's own build (asserts on)
    # aborts in curvelet::set_output for output type 2; the library writes it
    assert TABLE["_reference_build_type2_aborts"] is True


@pytest.mark.parametrize("nthreads", [1, 3, 8])
def test_thread_count_does_not_change_results(pkg, nthreads):
    e, H, W, _ = CASES["img_t_less_pgm"]
    p = TABLE["img_t_less_pgm"]["params"]
    a = run(pkg, e, H, W, p, nthreads=nthreads)
    b = run(pkg, e, H, W, p, nthreads=0)
    for t in (0, 1, 2):
        np.testing.assert_array_equal(a[t][0], b[t][0])
        np.testing.assert_array_equal(a[t][1].view(np.int64), b[t][1].view(np.int64))


def test_torch_and_row_major_inputs(pkg):
    import torch
    e, H, W, _ = CASES["scene_64x96_s1"]
    c1, i1 = pkg.form_curvelets(e, H, W)
    c2, i2 = pkg.form_curvelets(torch.from_numpy(e), H, W)
    c3, i3 = pkg.form_curvelets(np.asfortranarray(e), H, W)
    for c, i in ((c2, i2), (c3, i3)):
        np.testing.assert_array_equal(c, c1)
        np.testing.assert_array_equal(i, i1)


def test_straight_line_has_zero_curvature(pkg):
    # edgels every 0.5 px along a line at 30 degrees, tangent to it: every
    # curvelet's bundle contains k = 0 and the best fit is the anchor itself
    t = np.deg2rad(30.0)
    s = np.arange(0, 40, 0.5)
    e = np.stack([20 + s * np.cos(t), 10 + s * np.sin(t), np.full_like(s, t), np.full_like(s, 10.0)], 1)
    chain, info = pkg.form_curvelets(e, 64, 64)
    assert len(chain) > 0
    assert np.allclose(info[:, 7], 0.0, atol=1e-12)                 # curvature
    np.testing.assert_allclose(info[:, 4], e[chain[:, 0] - 1, 0])   # position = anchor
    np.testing.assert_allclose(info[:, 6], t)                       # orientation = tangent
    # every chain is consecutive along the line
    for row in chain[:, 1:]:
        ids = row[row > 0] - 1
        assert (np.abs(np.diff(ids)) >= 1).all() and np.ptp(ids) == len(ids) - 1


def test_circle_curvature_sign_and_size(pkg):
    # edgels on a circle of radius 12, tangents counter-clockwise: forward
    # curvelets estimate |k| near 1/12 (within the bundle grid's resolution)
    R = 12.0
    a = np.arange(0, 2 * np.pi, 0.5 / R)
    e = np.stack([32 + R * np.cos(a), 32 + R * np.sin(a), np.mod(a + np.pi / 2, 2 * np.pi), np.full_like(a, 5.0)],
                 1)
    chain, info = pkg.form_curvelets(e, 64, 64)
    fwd = info[:, 0] == 1.0
    assert fwd.any()
    k = info[fwd, 7]
    assert np.median(np.abs(k)) == pytest.approx(1 / R, rel=0.35)


@pytest.mark.parametrize("bad", [
    dict(nrows=0), dict(ncols=0), dict(dx=0.0), dict(dt=-0.1), dict(token_len=0.0), dict(rad=float("nan")),
])
def test_bad_arguments_raise(pkg, bad):
    e, H, W, _ = CASES["scene_64x96_s1"]
    kw = dict(nrows=H, ncols=W)
    kw.update(bad)
    nrows, ncols = kw.pop("nrows"), kw.pop("ncols")
    with pytest.raises(pkg.ToedError):
        pkg.CurveletFormation(e, nrows, ncols, **kw)


def test_output_before_execute_raises(pkg):
    e, H, W, _ = CASES["scene_64x96_s1"]
    cf = pkg.CurveletFormation(e, H, W)
    with pytest.raises(pkg.ToedError):
        cf.get_output_size(0)


def _ref_or_skip(oracle):
    if not oracle.curvelet_ref_available():
        pytest.skip("the compiled reference curvelet code (make -C oracle ref) is not present")


@pytest.mark.parametrize("seed", [11, 12])
def test_direct_against_compiled_reference_random(pkg, oracle, seed):
    """Fresh random lists (not in the golden table) and random parameters."""
    _ref_or_skip(oracle)
    rng = np.random.default_rng(seed)
    H, W = 80, 100
    n = 900
    e = np.stack([rng.uniform(0, W, n), rng.uniform(0, H, n), rng.uniform(0, 2 * np.pi, n),
                  rng.uniform(1, 9, n)], 1)
    p = dict(rad=float(rng.uniform(2.5, 4.5)), gap=float(rng.uniform(1.0, 2.5)), dx=float(rng.uniform(0.2, 0.6)),
             dt=float(np.deg2rad(rng.uniform(8, 20))), token_len=1.0, max_k=float(rng.uniform(0.2, 0.5)),
             style=int(seed % 4), group=int(rng.integers(3, 7)))
    _, ref = oracle.run_curvelet_ref(e, H, W, **p)
    out = run(pkg, e, H, W, p)
    for t in (0, 1, 2):
        np.testing.assert_array_equal(out[t][0], ref[t][0])
        np.testing.assert_array_equal(out[t][1].view(np.int64), ref[t][1].view(np.int64))


def test_direct_against_compiled_reference_oracle_list(pkg, oracle, scene):
    """The FP64 oracle's list of a fresh synthetic scene, the drivers' settings."""
    _ref_or_skip(oracle)
    img = scene.shapes_scene(21, 200, 240)
    e = oracle.detect(img)["edges"]
    _, ref = oracle.run_curvelet_ref(e, 200, 240, types="01")
    for t in (0, 1):
        c, i = pkg.form_curvelets(e, 200, 240, output_type=t)
        np.testing.assert_array_equal(c, ref[t][0])
        np.testing.assert_array_equal(i.view(np.int64), ref[t][1].view(np.int64))
