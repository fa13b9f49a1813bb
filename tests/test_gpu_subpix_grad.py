"""Differentiable subpixel correction on the GPU: dL/dmap from
SubpixBatch.backward (and the autograd wrapper) against FP64 central finite
differences of the oracle's frozen loss (oracle.SubpixFrozen: the same path
with every discrete decision of the forward frozen).  Tolerance: 1e-4
relative to the largest gradient entry of the map, per pixel, and 1e-4 of
sum |g_i d_i| on a random directional derivative sum g_i d_i over all pixels
(the sum cancels heavily, so its own size is no scale)."""
import glob
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

GOLD = os.path.join(os.path.dirname(__file__), "golden")
H_FD = 1e-6
TOL = 1e-4


def _fd(fr, E, G, d, h=H_FD):
    """Richardson-extrapolated central difference of the frozen loss along d
    (error O(h^4); weak NMS peaks make the loss strongly curved)."""
    c1 = (fr.loss(E + h * d, G) - fr.loss(E - h * d, G)) / (2 * h)
    c2 = (fr.loss(E + 0.5 * h * d, G) - fr.loss(E - 0.5 * h * d, G)) / h
    return (4 * c2 - c1) / 3


def _forward(pkg, detector, maps, thresh=0.1, sigma=2.0, nbr=2.0):
    dev = [torch.from_numpy(np.ascontiguousarray(E, np.float32)).cuda() for E in maps]
    b = pkg.SubpixBatch(detector, dev, None, thresh, sigma, nbr, with_index=True)
    b.set_grad(True)
    b.run()
    return b, dev


def _check_rows(outs, rows, label):
    """The frozen loss's rows at the forward point against the GPU forward's
    rows: the FP32 list's tolerances (position 1e-4 px, orientation 1e-4 rad
    modulo 2 pi, strength 1e-5 relative)."""
    rows = np.asarray(rows, np.float64)
    assert outs.shape == rows.shape, (label, outs.shape, rows.shape)
    dxy = np.abs(outs[:, :2] - rows[:, :2]).max()
    dth = np.abs(np.angle(np.exp(1j * (outs[:, 2] - rows[:, 2])))).max()
    dst = (np.abs(outs[:, 3] - rows[:, 3]) / np.maximum(np.abs(outs[:, 3]), 1e-12)).max()
    print("%s: frozen rows vs GPU rows: |dxy| %.2e px, |dtheta| %.2e rad, strength %.2e rel" % (label, dxy, dth, dst),
          flush=True)
    assert dxy <= 1e-4 and dth <= 1e-4 and dst <= 1e-5


def _check_map(oracle, E, idx, G, g, sigma, rng, label, rows=None, theta=None):
    fr = oracle.SubpixFrozen(E, idx, sigma, theta)
    L0, outs = fr.loss(E, G, outputs=True)
    if rows is not None:
        _check_rows(outs, rows, label)
    scale = np.abs(g).max()
    assert scale > 0
    # random directional derivative over every pixel
    d = rng.standard_normal(E.shape)
    fd = _fd(fr, E, G, d)
    an = float((g * d).sum())
    l1 = float(np.abs(g * d).sum())
    print("%s: N=%d max|g|=%.4g dir fd=%.8g gpu=%.8g |diff|/sum|g d|=%.2e" % (label, len(idx), scale, fd, an,
                                                                             abs(fd - an) / l1), flush=True)
    # single pixels: the largest entries and random pixels near edgels
    flat = np.abs(g).reshape(-1)
    top = np.argsort(flat)[-24:]
    nz = np.flatnonzero(flat > 1e-3 * scale)
    pick = np.unique(np.concatenate([top, rng.choice(nz, size=min(24, len(nz)), replace=False)]))
    worst, wself = 0.0, 0.0
    for p in pick:
        e = np.zeros(E.size)
        e[p] = 1.0
        fdp = _fd(fr, E, G, e.reshape(E.shape))
        err = abs(fdp - g.reshape(-1)[p])
        worst = max(worst, err / scale)
        wself = max(wself, err / max(abs(fdp), 1e-3 * scale))
    print("%s: %d pixels, worst |fd - gpu| / max|g| = %.2e, / |fd| = %.2e" % (label, len(pick), worst, wself),
          flush=True)
    assert worst <= TOL
    assert abs(fd - an) <= TOL * l1


def test_subpix_grad_matches_fd(pkg, detector, oracle):
    z = np.load(os.path.join(GOLD, "subpix_ref_bezier_96x128_s1.npz"))
    E = z["E"].astype(np.float64)
    b, dev = _forward(pkg, detector, [z["E"]])
    e = b.results()[0]
    idx = b.index_results()[0].cpu().numpy().astype(np.int64)
    ref = oracle.subpix(E, None, 0.1, 2.0, 2.0)
    assert np.array_equal(idx, ref["idx"])
    rng = np.random.default_rng(5)
    G = rng.standard_normal((len(idx), 4)).astype(np.float32)
    gmap = torch.zeros_like(dev[0])
    b.backward([torch.from_numpy(G).cuda()], [gmap])
    g = gmap.cpu().numpy().astype(np.float64)
    _check_map(oracle, E, idx, G.astype(np.float64), g, 2.0, rng, "96x128", rows=e.cpu().numpy())
    # the backward accumulates: a second call doubles the gradient
    b.backward([torch.from_numpy(G).cuda()], [gmap])
    np.testing.assert_allclose(gmap.cpu().numpy(), 2 * g, rtol=1e-5, atol=1e-6 * np.abs(g).max())
    b.close()


def test_subpix_grad_given_orientation(pkg, detector, oracle):
    """A map with a given orientation map (set_orientation): the direction is a
    constant input of the adjoint; FD of the frozen loss with the same FP32
    orientation values."""
    z = np.load(os.path.join(GOLD, "subpix_ref_bezier_128x160_s4_orient.npz"))
    thr, sig, nbr = (float(v) for v in z["params"])
    E32 = np.ascontiguousarray(z["E"], np.float32)
    th32 = np.ascontiguousarray(z["theta"], np.float32)
    E, th = E32.astype(np.float64), th32.astype(np.float64)
    dev = [torch.from_numpy(E32).cuda()]
    b = pkg.SubpixBatch(detector, dev, [torch.from_numpy(th32).cuda()], thr, sig, nbr, with_index=True)
    b.set_grad(True)
    b.run()
    e = b.results()[0].cpu().numpy()
    idx = b.index_results()[0].cpu().numpy().astype(np.int64)
    ref = oracle.subpix(E, th, thr, sig, nbr)
    assert np.array_equal(idx, ref["idx"]) and len(idx) > 100
    rng = np.random.default_rng(13)
    G = rng.standard_normal((len(idx), 4)).astype(np.float32)
    gmap = torch.zeros_like(dev[0])
    b.backward([torch.from_numpy(G).cuda()], [gmap])
    g = gmap.cpu().numpy().astype(np.float64)
    _check_map(oracle, E, idx, G.astype(np.float64), g, sig, rng, "given orientation 128x160", rows=e, theta=th)
    b.close()


def test_subpix_grad_batch_and_autograd(pkg, detector, oracle, scene):
    maps = [scene.bezier_map(71, 80, 120, 8), scene.bezier_map(72, 112, 96, 8)]
    B = len(maps)
    b, dev = _forward(pkg, detector, maps, 0.1, 1.5, 2.0)
    res, idx = b.results(), b.index_results()
    rng = np.random.default_rng(9)
    Gs = [rng.standard_normal((len(r), 4)).astype(np.float32) for r in res]
    gmaps = [torch.zeros_like(d) for d in dev]
    b.backward([torch.from_numpy(G).cuda() for G in Gs], gmaps)
    for m in range(B):
        _check_map(oracle, maps[m].astype(np.float64), idx[m].cpu().numpy().astype(np.int64),
                   Gs[m].astype(np.float64), gmaps[m].cpu().numpy().astype(np.float64), 1.5, rng, "map %d" % m,
                   rows=res[m].cpu().numpy())
    b.close()
    # the autograd wrapper on a [B, H, W] stack gives the same as the plan API
    E = scene.bezier_map(73, 96, 128, 10)
    x = torch.from_numpy(np.stack([E, E[::-1].copy()])).cuda().requires_grad_(True)
    w = torch.tensor([0.3, -1.2, 0.7, 2.0], device="cuda")
    edges, off = pkg.subpix_differentiable(x, 0.1, 2.0, 2.0, detector=detector)
    (edges * w).sum().backward()
    gx = x.grad.cpu().numpy().astype(np.float64)
    for m in range(2):
        Em = x[m].detach().cpu().numpy().astype(np.float64)
        rows = edges[int(off[m]):int(off[m + 1])]
        ref = oracle.subpix(Em, None, 0.1, 2.0, 2.0)
        assert len(rows) == ref["n"]
        fr = oracle.SubpixFrozen(Em, ref["idx"], 2.0)
        d = np.random.default_rng(m).standard_normal(Em.shape)
        wn = w.cpu().numpy().astype(np.float64)
        _check_rows(fr.loss(Em, wn, outputs=True)[1], rows.detach().cpu().numpy(), "autograd map %d" % m)
        fd = _fd(fr, Em, wn, d)
        an = float((gx[m] * d).sum())
        l1 = float(np.abs(gx[m] * d).sum())
        print("autograd map %d: fd %.8g gpu %.8g |diff|/sum|g d| %.2e" % (m, fd, an, abs(fd - an) / l1), flush=True)
        assert abs(fd - an) <= TOL * l1


def test_subpix_plan_rerun_on_sparser_map(pkg, detector, oracle, scene):
    """A plan's work planes and masks carry nothing from one run to the next:
    the same plan run on a dense map and then on a sparse one (most of whose
    tiles have no edgel, and whose dense backward skips them) gives the
    sparse map's own edgels and gradient (against the oracle), and the
    forward agrees with a fresh plan's."""
    dense = scene.bezier_map(81, 320, 384, 40)
    sparse = scene.bezier_map(82, 320, 384, 2)
    b, dev = _forward(pkg, detector, [dense])
    b.results()
    dev[0].copy_(torch.from_numpy(sparse))
    b.run()
    e = b.results()[0].cpu().numpy()
    idx = b.index_results()[0].cpu().numpy().astype(np.int64)
    ref = oracle.subpix(sparse.astype(np.float64), None, 0.1, 2.0, 2.0)
    assert np.array_equal(idx, ref["idx"]) and len(idx) > 20
    fresh, fdev = _forward(pkg, detector, [sparse])
    assert np.array_equal(fresh.results()[0].cpu().numpy(), e)
    fresh.close()
    rng = np.random.default_rng(11)
    G = rng.standard_normal((len(idx), 4)).astype(np.float32)
    gmap = torch.zeros_like(dev[0])
    b.backward([torch.from_numpy(G).cuda()], [gmap])
    g = gmap.cpu().numpy().astype(np.float64)
    assert np.isfinite(g).all()
    _check_map(oracle, sparse.astype(np.float64), idx, G.astype(np.float64), g, 2.0, rng, "rerun 320x384", rows=e)
    b.close()


def test_subpix_grad_full_size_maps(pkg, detector, oracle, scene):
    """The differentiable correction at the workload's map size: two Bezier
    maps of 2048 x 2048 (40 curves, the benchmark's seeds), thresh 0.1, sigma 2,
    orientation estimated.  The plan API's gradient is linear in the upstream
    rows, the autograd wrapper gives the plan API's rows and gradient, the
    frozen FP64 loss reproduces map 0's rows, and a random directional
    derivative of it matches the GPU gradient to 1e-4 of sum |g d|."""
    S, B, cap = 2048, 2, 2048 * 2048 // 8
    maps = [scene.bezier_map(5000 + m, S, S, 40) for m in range(B)]
    x = torch.from_numpy(np.stack(maps)).cuda()
    b = pkg.SubpixBatch(detector, [x[m] for m in range(B)], None, 0.1, 2.0, 2.0, capacities=[cap] * B,
                        with_index=True)
    b.set_grad(True)
    b.run()
    res = [r.cpu().numpy() for r in b.results()]
    idx = [i.cpu().numpy().astype(np.int64) for i in b.index_results()]
    assert all(len(r) > 50000 for r in res)
    rng = np.random.default_rng(21)
    G1 = [rng.standard_normal((len(r), 4)).astype(np.float32) for r in res]
    G2 = [rng.standard_normal((len(r), 4)).astype(np.float32) for r in res]

    def grad(Gs):
        gm = [torch.zeros((S, S), dtype=torch.float32, device="cuda") for _ in range(B)]
        b.backward([torch.from_numpy(np.ascontiguousarray(G)).cuda() for G in Gs], gm)
        return [g.cpu().numpy().astype(np.float64) for g in gm]

    g1, g2 = grad(G1), grad(G2)
    g12 = grad([2 * a + 0.5 * c for a, c in zip(G1, G2)])
    for m in range(B):
        scale = np.abs(g12[m]).max()
        assert scale > 0
        assert np.abs(g12[m] - (2 * g1[m] + 0.5 * g2[m])).max() <= 1e-5 * scale
    b.close()
    # the autograd wrapper on the [B, S, S] stack
    xg = x.clone().requires_grad_(True)
    e, off = pkg.subpix_differentiable(xg, 0.1, 2.0, 2.0, detector=detector, capacity=cap)
    assert np.array_equal(e.detach().cpu().numpy(), np.concatenate(res))
    (e * torch.from_numpy(np.concatenate(G1)).cuda()).sum().backward()
    for m in range(B):
        ga = xg.grad[m].cpu().numpy().astype(np.float64)
        assert np.abs(ga - g1[m]).max() <= 1e-5 * np.abs(g1[m]).max()
    # FP64 frozen loss of map 0: its rows, then one random directional derivative
    E0 = maps[0].astype(np.float64)
    fr = oracle.SubpixFrozen(E0, idx[0], 2.0)
    _check_rows(fr.loss(E0, G1[0].astype(np.float64), outputs=True)[1], res[0], "2048 map 0")
    d = rng.standard_normal(E0.shape)
    fd = _fd(fr, E0, G1[0].astype(np.float64), d)
    an = float((g1[0] * d).sum())
    l1 = float(np.abs(g1[0] * d).sum())
    print("2048 map 0: %d edgels, directional fd %.8g gpu %.8g |diff|/sum|g d| %.2e"
          % (len(idx[0]), fd, an, abs(fd - an) / l1), flush=True)
    assert abs(fd - an) <= TOL * l1


def test_subpix_autograd_plan_reuse_and_rerun(pkg, detector, scene):
    """The subpix autograd plan is keyed by shape and parameters: a new maps
    tensor each step reuses it, and a backward whose plan has run again since
    its forward (a second forward in between) re-runs that forward's maps
    first and gives the same gradient as right after it."""
    sp = __import__(pkg.__name__ + ".subpix", fromlist=["_SP_PLANS"])
    base = np.stack([scene.bezier_map(90 + m, 96, 128, 8) for m in range(2)])
    w = torch.tensor([0.5, -1.1, 0.8, 1.7], device="cuda")
    x1 = torch.from_numpy(base.copy()).cuda().requires_grad_(True)
    e1, _ = pkg.subpix_differentiable(x1, 0.1, 2.0, 2.0, detector=detector)
    (e1 * w).sum().backward()
    g_ref = x1.grad.clone()
    plan = next(reversed(sp._SP_PLANS.values()))
    x2 = torch.from_numpy(base.copy()).cuda().requires_grad_(True)     # a new tensor, same data
    e2, _ = pkg.subpix_differentiable(x2, 0.1, 2.0, 2.0, detector=detector)
    x3 = torch.from_numpy(base[::-1].copy()).cuda()                   # other data through the same plan
    pkg.subpix_differentiable(x3, 0.1, 2.0, 2.0, detector=detector)
    assert next(reversed(sp._SP_PLANS.values())) is plan and len(sp._SP_PLANS) <= 2
    (e2 * w).sum().backward()                                            # re-runs x2's forward first
    assert torch.equal(e1, e2)
    assert (x2.grad - g_ref).abs().max().item() <= 1e-5 * g_ref.abs().max().item()
