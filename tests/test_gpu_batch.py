"""Batched form: one plan over images of uneven sizes and sigma, plain and
persistent tile kernels, and hipGraph replays, each checked edgel for edgel
against single-image detection and the oracle."""
import numpy as np
import pytest

from parity import assert_parity, compare

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

SHAPES = [(480, 640), (120, 200), (64, 900), (333, 257), (20, 20), (481, 321)]
SIGMAS = [1.0, 1.5, 2.0, 2.0, 1.5, 1.0]


@pytest.fixture(scope="module")
def batch_inputs(scene):
    imgs = [scene.shapes_scene(500 + k, H, W) for k, (H, W) in enumerate(SHAPES)]
    return imgs, [torch.from_numpy(im).cuda() for im in imgs]


def _singles(detector, dev_imgs):
    out = []
    for t, s in zip(dev_imgs, SIGMAS):
        e, i = detector.detect(t, sigma=s, with_index=True)
        out.append((e.cpu().numpy(), i.cpu().numpy()))
    return out


@pytest.mark.parametrize("persistent", [False, True])
def test_batch_equals_singles(pkg, detector, batch_inputs, persistent):
    imgs, dev = batch_inputs
    singles = _singles(detector, dev)
    b = pkg.BatchDetector(detector, dev, SIGMAS, persistent=persistent, with_index=True)
    b.run()
    res = b.results()
    idx = b.index_results()
    for m in range(len(dev)):
        assert np.array_equal(idx[m].cpu().numpy(), singles[m][1]), m
        assert np.array_equal(res[m].cpu().numpy(), singles[m][0]), m
    b.close()


def test_batch_graph_replays(pkg, detector, batch_inputs, oracle):
    imgs, dev = batch_inputs
    b = pkg.BatchDetector(detector, dev, SIGMAS, persistent=True, with_index=True)
    b.run()
    first = [r.cpu().numpy().copy() for r in b.results()]
    for rep in range(4):
        b.edges.fill_(-1.0)   # eager work between replays
        b.counts.zero_()
        torch.cuda.synchronize()
        b.run(graph=True)
        torch.cuda.synchronize()
        now = [r.cpu().numpy() for r in b.results()]
        for m in range(len(dev)):
            assert np.array_equal(now[m], first[m]), (rep, m)
    # and the batch agrees with the oracle
    idx = [i.cpu().numpy() for i in b.index_results()]
    for m in (0, 3):
        ref = oracle.detect(imgs[m].astype(np.float64), SIGMAS[m], margins=True)
        assert_parity(compare(ref, first[m], idx[m]))
    b.close()


def test_autotune_times_every_config_and_keeps_results(pkg, batch_inputs):
    """The autotuner times tile rows 8 and 16, each with one workgroup per
    tile and with persistent grids of 1 .. (what fits) workgroups per CU,
    keeps the fastest, and a plan built with its choice (and with every
    persistent grid size) gives the default plan's edgels."""
    imgs, dev = batch_inputs
    det = pkg.Detector(0)
    ref = pkg.BatchDetector(det, dev, SIGMAS)
    ref.run()
    base = [r.cpu().numpy().copy() for r in ref.results()]
    ref.close()
    tune = det.autotune_batch(dev, SIGMAS, reps=2)
    res, best = tune["results"], tune["best"]
    print("autotune:", best, "of", len(res))
    assert {r["tile_rows"] for r in res} == {8, 16}
    assert {(r["tile_rows"], r["persistent"]) for r in res} == {(8, False), (8, True), (16, False), (16, True)}
    assert all(r["ms"] > 0 for r in res) and all((r["wg_per_cu"] > 0) == r["persistent"] for r in res)
    assert best["ms"] == min(r["ms"] for r in res)
    grids = sorted({r["wg_per_cu"] for r in res if r["persistent"]})
    assert grids[0] == 1 and grids == list(range(1, grids[-1] + 1))
    for w in [best["wg_per_cu"]] + grids:
        b = pkg.BatchDetector(det, dev, SIGMAS, persistent=best["persistent"] if w == best["wg_per_cu"] else True,
                              wg_per_cu=w)
        b.run(graph=True)
        out = [r.cpu().numpy() for r in b.results()]
        for m in range(len(dev)):
            assert np.array_equal(out[m], base[m]), (w, m)
        b.close()
    det.close()


def test_batch_float_images(pkg, detector, batch_inputs):
    imgs, dev = batch_inputs
    devf = [t.float() for t in dev[:3]]
    b = pkg.BatchDetector(detector, devf, SIGMAS[:3])
    b.run()
    res = b.results()
    b8 = pkg.BatchDetector(detector, dev[:3], SIGMAS[:3])
    b8.run()
    res8 = b8.results()
    for m in range(3):
        assert torch.equal(res[m], res8[m])
    b.close()
    b8.close()


def test_full_batch_200_mixed_equals_singles(pkg, detector, scene):
    # the batched workload at its full size: 200 scenes, sizes drawn from
    # {480x640, 480x752, 720x1280, 1080x1920, 2160x3840}, sigma from {1, 1.5, 2};
    # the captured persistent plan must give each image's single-image list
    sizes = [(480, 640), (480, 752), (720, 1280), (1080, 1920), (2160, 3840)]
    rng = np.random.default_rng(2024)
    shapes = [sizes[k] for k in rng.integers(0, len(sizes), 200)]
    sigmas = [[1.0, 1.5, 2.0][k] for k in rng.integers(0, 3, 200)]
    dev = [torch.from_numpy(scene.shapes_scene(2024 * 1000 + m, H, W)).cuda() for m, (H, W) in enumerate(shapes)]
    b = pkg.BatchDetector(detector, dev, sigmas, persistent=True, with_index=True)
    b.run(graph=True)
    res = [r.cpu().numpy() for r in b.results()]
    idx = [i.cpu().numpy() for i in b.index_results()]
    b.close()
    total = 0
    for m, (t, s) in enumerate(zip(dev, sigmas)):
        e, i = detector.detect(t, sigma=s, with_index=True)
        assert np.array_equal(idx[m], i.cpu().numpy()), m
        assert np.array_equal(res[m], e.cpu().numpy()), m
        total += len(res[m])
    assert total > 1000000
