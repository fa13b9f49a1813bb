"""Subpixel third-order correction path on the GPU (SubpixBatch over the C
ABI) against the FP64 oracle, which itself equals the reference's own
subpix_to_correction.cpp bit for bit on the golden maps."""
import glob
import os

import numpy as np
import pytest

from parity import assert_subpix_parity, compare_subpix

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

GOLDEN = sorted(glob.glob(os.path.join(os.path.dirname(__file__), "golden", "subpix_ref_*.npz")))


def _run(pkg, detector, maps, orients=None, thresh=0.1, sigma=2.0, nbr=2.0, with_maps=False, caps=None, interp_n=0):
    dev = [torch.from_numpy(np.ascontiguousarray(E, np.float32)).cuda() for E in maps]
    ors = None
    if orients is not None:
        ors = [None if o is None else torch.from_numpy(np.ascontiguousarray(o, np.float32)).cuda() for o in orients]
    b = pkg.SubpixBatch(detector, dev, ors, thresh, sigma, nbr, capacities=caps, with_index=True,
                        with_maps=with_maps, interp_n=interp_n)
    b.run()
    return b


@pytest.mark.parametrize("path", GOLDEN, ids=[os.path.basename(p)[11:-4] for p in GOLDEN])
def test_subpix_golden(pkg, detector, oracle, path):
    """Every golden map, including those of the class constructed with
    n = interp_n > 0: the GPU works on float maps (the upsampled work grid is
    stored as float), so the oracle gets the same rounded values and the edgel
    sets must be identical."""
    z = np.load(path)
    thr, sig, nbr = (float(v) for v in z["params"])
    theta = z["theta"] if "theta" in z.files else None
    interp_n = int(z["interp_n"]) if "interp_n" in z.files else 0
    # the GPU takes float32 orientation maps: the oracle gets the same rounded values
    th32 = None if theta is None else theta.astype(np.float32)
    ref = oracle.subpix(z["E"].astype(np.float64), None if th32 is None else th32.astype(np.float64), thr, sig,
                        nbr, maps=True, interp_n=interp_n, work_f32=interp_n > 0)
    if th32 is None and interp_n == 0:
        assert np.array_equal(ref["edges"], z["edges"])
    print("map %s sigma %g thresh %g r %g interp_n %d" % (z["E"].shape, sig, thr, nbr, interp_n), flush=True)
    b = _run(pkg, detector, [z["E"]], None if th32 is None else [th32], thr, sig, nbr, with_maps=True,
             interp_n=interp_n)
    e = b.results()[0].cpu().numpy()
    i = b.index_results()[0].cpu().numpy()
    rep = compare_subpix(ref, e, i)
    print(rep)
    assert_subpix_parity(rep)
    te = b.to_edgemap[0].cpu().numpy()
    to = b.to_orientation[0].cpu().numpy()
    assert np.array_equal(te != 0, ref["TO_E"] != 0)
    np.testing.assert_allclose(te, ref["TO_E"], rtol=1e-4, atol=1e-7)
    nz = ref["TO_E"] != 0
    assert np.abs(np.angle(np.exp(1j * (to[nz] - ref["TO_O"][nz])))).max() < 1e-3
    b.close()


def test_subpix_mixed_batch(pkg, detector, oracle, scene):
    shapes = [(96, 128), (200, 300), (64, 517), (333, 257), (12, 12), (9, 9)]
    maps = [scene.bezier_map(40 + k, H, W, 12) for k, (H, W) in enumerate(shapes)]
    rng = np.random.default_rng(7)
    orients = [None, rng.uniform(-3.0, 3.0, shapes[1]).astype(np.float32), None, None, None, None]
    thr = [0.1, 0.2, 0.15, 0.1, 0.1, 0.1]
    print("mixed batch", shapes, flush=True)
    b = _run(pkg, detector, maps, orients, thr, 1.5, 2.0)
    res, idx = b.results(), b.index_results()
    for m in range(len(maps)):
        ref = oracle.subpix(maps[m].astype(np.float64),
                            None if orients[m] is None else orients[m].astype(np.float64), thr[m], 1.5, 2.0)
        rep = compare_subpix(ref, res[m].cpu().numpy(), idx[m].cpu().numpy())
        print(m, rep)
        assert_subpix_parity(rep)
    assert len(res[-1]) == 0          # 9 x 9: NMS rows [4, 4) are empty
    b.close()


@pytest.mark.parametrize("sigma", [1.0, 3.0])
def test_subpix_sigma(pkg, detector, oracle, scene, sigma):
    E = scene.bezier_map(11, 256, 320, 16)
    print("sigma", sigma, flush=True)
    b = _run(pkg, detector, [E], None, 0.1, sigma, 2.0)
    ref = oracle.subpix(E.astype(np.float64), None, 0.1, sigma, 2.0)
    rep = compare_subpix(ref, b.results()[0].cpu().numpy(), b.index_results()[0].cpu().numpy())
    print(rep)
    assert_subpix_parity(rep)
    b.close()


def test_subpix_full_size_map(pkg, detector, oracle, scene):
    # one map of the benchmark shape: 2048 x 2048, 40 curves, thresh 0.1, sigma 2
    E = scene.bezier_map(2024, 2048, 2048, 40)
    print("2048 x 2048", flush=True)
    b = _run(pkg, detector, [E], None, 0.1, 2.0, 2.0)
    ref = oracle.subpix(E.astype(np.float64), None, 0.1, 2.0, 2.0, nthreads=8)
    rep = compare_subpix(ref, b.results()[0].cpu().numpy(), b.index_results()[0].cpu().numpy())
    print(rep)
    assert rep["n_oracle"] > 50000
    assert_subpix_parity(rep)
    # repeated runs are bit-identical
    e1 = b.results()[0].clone()
    b.run()
    assert torch.equal(e1, b.results()[0])
    b.close()


def test_subpix_capacity_flag(pkg, detector, scene):
    E = scene.bezier_map(5, 128, 128, 12)
    b = _run(pkg, detector, [E], caps=[10])
    with pytest.raises(pkg.ToedError):
        b.results()
    assert int(b.counts[0]) > 10          # the count stays exact
    b.close()


def test_subpix_empty_and_flat_maps(pkg, detector):
    b = _run(pkg, detector, [np.zeros((64, 64), np.float32), np.full((40, 50), 0.5, np.float32)])
    assert [len(r) for r in b.results()] == [0, 0]
    b.close()


def test_subpix_class_mirror(pkg, detector, oracle, scene):
    E = scene.bezier_map(3, 120, 150, 10)
    c = pkg.SubpixelTOCorrectionGPU(120, 150, 0.1, 2.0, 0)
    c.set_edgemap(E)
    c.estimate_orientation()
    out = np.zeros((120 * 150, 4), np.float32)
    n = c.run(out)
    ref = oracle.subpix(E.astype(np.float64), None, 0.1, 2.0, 2.0, maps=True)
    assert n == ref["n"] == c.edge_pt_list_idx
    np.testing.assert_allclose(out[:n, :2], ref["edges"][:, :2], atol=1e-3)
    np.testing.assert_allclose(c.TO_edgemap, ref["TO_E"], rtol=1e-4, atol=1e-7)
    c.close()


def _check_interp(ref, b, m, shape):
    rep = compare_subpix(ref, b.results()[m].cpu().numpy(), b.index_results()[m].cpu().numpy())
    print(m, shape, rep, flush=True)
    assert_subpix_parity(rep)
    te = b.to_edgemap[m].cpu().numpy()
    assert te.shape == shape
    assert np.array_equal(te != 0, ref["TO_E"] != 0)
    np.testing.assert_allclose(te, ref["TO_E"], rtol=1e-4, atol=1e-7)


@pytest.mark.parametrize("interp_n", [1, 2])
def test_subpix_interp_mixed_batch(pkg, detector, oracle, scene, interp_n):
    """interp_n > 0 over maps of mixed and ragged sizes (the work grid
    (H - 1) 2^n + 1 is never a whole number of tiles), one with a given
    orientation: identical edgel sets to the oracle on the same float work
    maps, x and y in map pixels, TO maps of the map size.  At n = 2 the
    isolation reach on the work grid is floor(2 * 4 + 2 sqrt 2) = 10 pixels
    (the wide isolation kernel)."""
    shapes = [(96, 128), (70, 133), (33, 300), (9, 9), (2, 2)]
    maps = [scene.bezier_map(60 + k, H, W, 8) for k, (H, W) in enumerate(shapes)]
    rng = np.random.default_rng(interp_n)
    orients = [None, rng.uniform(-3.0, 3.0, shapes[1]).astype(np.float32), None, None, None]
    print("interp_n", interp_n, shapes, flush=True)
    b = _run(pkg, detector, maps, orients, 0.1, 2.0, 2.0, with_maps=True, interp_n=interp_n)
    for m in range(len(maps)):
        ref = oracle.subpix(maps[m].astype(np.float64), None if orients[m] is None else orients[m].astype(np.float64),
                            0.1, 2.0, 2.0, maps=True, interp_n=interp_n, work_f32=True)
        _check_interp(ref, b, m, shapes[m])
    assert len(b.results()[0]) > 500
    # re-pointed maps (toed_subpix_set_maps re-resizes them at the next run)
    b.set_maps([torch.from_numpy(np.ascontiguousarray(maps[0][::-1], np.float32)).cuda()] +
               [torch.from_numpy(np.ascontiguousarray(E, np.float32)).cuda() for E in maps[1:]])
    b.run()
    ref = oracle.subpix(maps[0][::-1].astype(np.float64), None, 0.1, 2.0, 2.0, maps=True, interp_n=interp_n,
                        work_f32=True)
    _check_interp(ref, b, 0, shapes[0])
    b.close()


def test_subpix_interp_limits(pkg, detector, scene):
    E = torch.from_numpy(scene.bezier_map(1, 40, 40, 4)).cuda()
    # radius 3 on a 2^2 work grid reaches floor(12 + 2 sqrt 2) = 14 > 12 work pixels
    with pytest.raises(pkg.ToedError):
        pkg.SubpixBatch(detector, [E], None, 0.1, 2.0, 3.0, interp_n=2)
    with pytest.raises(pkg.ToedError):
        pkg.SubpixBatch(detector, [E], None, 0.1, 2.0, 2.0, interp_n=5)
    b = pkg.SubpixBatch(detector, [E], None, 0.1, 2.0, 2.0, interp_n=1)
    with pytest.raises(pkg.ToedError):
        b.set_grad(True)          # the adjoint covers interp_n = 0
    b.close()


def test_subpix_class_mirror_interp(pkg, detector, oracle, scene):
    E = scene.bezier_map(4, 90, 110, 8)
    c = pkg.SubpixelTOCorrectionGPU(90, 110, 0.1, 2.0, 0, 1)
    c.set_edgemap(E)
    c.estimate_orientation()
    wh, ww = 179, 219
    out = np.zeros((wh * ww, 4), np.float32)
    n = c.run(out)
    ref = oracle.subpix(E.astype(np.float64), None, 0.1, 2.0, 2.0, maps=True, interp_n=1, work_f32=True)
    assert n == ref["n"] == c.edge_pt_list_idx
    np.testing.assert_allclose(out[:n, :2], ref["edges"][:, :2], atol=1e-3)
    np.testing.assert_allclose(c.TO_edgemap, ref["TO_E"], rtol=1e-4, atol=1e-7)
    c.close()


def test_subpix_interp_large_map(pkg, detector, oracle, scene):
    """interp_n = 1 on a 768 x 1024 map (work grid 1535 x 2047, 3.1 M work pixels): identical
    edgel set to the oracle on the same float work map, TO maps of the map size, and
    repeated runs bit-identical."""
    E = scene.bezier_map(77, 768, 1024, 24)
    print("768 x 1024, interp_n 1", flush=True)
    b = _run(pkg, detector, [E], None, 0.1, 2.0, 2.0, with_maps=True, interp_n=1)
    ref = oracle.subpix(E.astype(np.float64), None, 0.1, 2.0, 2.0, nthreads=8, maps=True, interp_n=1, work_f32=True)
    _check_interp(ref, b, 0, E.shape)
    assert ref["n"] > 20000
    e1 = b.results()[0].clone()
    b.run()
    assert torch.equal(e1, b.results()[0])
    b.close()
