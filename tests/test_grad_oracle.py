"""The FP64 gradient oracle of the differentiable form, pinned against
central finite differences of the frozen-edgel-set loss (CPU only)."""
import numpy as np
import pytest


def _setup(oracle, scene, seed=6, H=48, W=64, sigma=1.5):
    img = scene.shapes_scene(seed, H, W).astype(np.float64)
    det = oracle.detect(img, sigma)
    eidx = det["eidx"]
    rng = np.random.default_rng(seed)
    w = rng.normal(size=(len(eidx), 4))
    assert len(eidx) > 100
    L, oc = oracle.frozen_loss(img, sigma, eidx, w)
    return img, eidx, w, oc


def test_frozen_loss_reproduces_detection(oracle, scene):
    """With the octants decided on the image, the frozen-set outputs are the
    detector's own edgel rows (cpu_toed.cpp:516-529)."""
    img = scene.shapes_scene(4, 60, 80).astype(np.float64)
    for sigma in (1.0, 2.0):
        det = oracle.detect(img, sigma)
        L, oc, outs = oracle.frozen_loss(img, sigma, det["eidx"], np.zeros(4), outputs=True)
        e = det["edges"]
        assert np.abs(outs[:, :2] - e[:, :2]).max() < 1e-9
        assert np.abs(outs[:, 3] - e[:, 3]).max() < 1e-9
        dth = np.abs(outs[:, 2] - e[:, 2])
        assert np.minimum(dth, np.abs(dth - 2 * np.pi)).max() < 1e-9
        assert np.all((oc >= 1) & (oc <= 8))


def test_grad_oracle_matches_finite_differences(oracle, scene):
    img, eidx, w, oc = _setup(oracle, scene)
    g, gs = oracle.frozen_grad(img, 1.5, eidx, w, oc)
    assert np.count_nonzero(g) > 100
    rng = np.random.default_rng(0)
    cand = np.flatnonzero(np.abs(g) > 1e-3 * np.abs(g).max())
    pix = rng.choice(cand, size=40, replace=False)
    h = 1e-3
    for k in pix:
        i, j = divmod(int(k), img.shape[1])
        ip, im = img.copy(), img.copy()
        ip[i, j] += h
        im[i, j] -= h
        fd = (oracle.frozen_loss(ip, 1.5, eidx, w, octants=oc)[0] -
              oracle.frozen_loss(im, 1.5, eidx, w, octants=oc)[0]) / (2 * h)
        assert fd == pytest.approx(g[i, j], rel=1e-5, abs=1e-7 * np.abs(g).max())
    # sigma, with the window radii frozen
    hs = 1e-5
    R0, R1 = oracle.radius(1.5)
    fds = (oracle.frozen_loss(img, 1.5 + hs, eidx, w, R0, R1, octants=oc)[0] -
           oracle.frozen_loss(img, 1.5 - hs, eidx, w, R0, R1, octants=oc)[0]) / (2 * hs)
    assert fds == pytest.approx(gs, rel=1e-5)


def test_grad_is_linear_in_weights(oracle, scene):
    img, eidx, w, oc = _setup(oracle, scene, seed=9)
    g1, s1 = oracle.frozen_grad(img, 1.5, eidx, w, oc)
    g2, s2 = oracle.frozen_grad(img, 1.5, eidx, 2.5 * w, oc)
    np.testing.assert_allclose(g2, 2.5 * g1, rtol=1e-6, atol=1e-9 * np.abs(g1).max())
    assert s2 == pytest.approx(2.5 * s1, rel=1e-6)
    # a pure x-shift weight on the x column only: gradient of sum(x)
    gx, _ = oracle.frozen_grad(img, 1.5, eidx, np.array([1.0, 0, 0, 0]), oc)
    gy, _ = oracle.frozen_grad(img, 1.5, eidx, np.array([0, 1.0, 0, 0]), oc)
    gsum, _ = oracle.frozen_grad(img, 1.5, eidx, np.array([1.0, 1.0, 0, 0]), oc)
    np.testing.assert_allclose(gsum, gx + gy, rtol=1e-6, atol=1e-9 * np.abs(gsum).max())
