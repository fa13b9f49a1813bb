"""Differentiable form on the GPU: gradients of the edgel loss
sum(a x + b y + c theta + d strength) w.r.t. the pixels and sigma, with the
edgel set frozen, against (1) FP64 central finite differences of the oracle's
frozen-set loss and (2) the oracle's independent FP64 gradient."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def _forward_backward(pkg, imgs_np, sigma, weights):
    imgs = torch.from_numpy(np.stack(imgs_np).astype(np.float32)).cuda().requires_grad_(True)
    sig = torch.tensor(float(sigma), dtype=torch.float64, requires_grad=True)
    edges, offsets = pkg.detect_differentiable(imgs, sig)
    w = torch.tensor(weights, dtype=torch.float32, device="cuda")
    loss = (edges * w).sum()
    loss.backward()
    torch.cuda.synchronize()
    return edges.detach().cpu().numpy(), offsets.numpy(), imgs.grad.cpu().numpy(), float(sig.grad)


@pytest.mark.parametrize("sigma", [1.0, 1.5, 2.0])
def test_grad_matches_oracle_and_fd(pkg, detector, scene, oracle, sigma):
    rng = np.random.default_rng(7)
    w = rng.normal(size=4)
    imgs = [scene.shapes_scene(60 + k, 96, 128) for k in range(2)]
    edges, off, g, gs = _forward_backward(pkg, imgs, sigma, w)
    gs_ref = 0.0
    for b, im in enumerate(imgs):
        n = off[b + 1] - off[b]
        assert n > 50
        # frozen set = the GPU's edgels (interp indices from the forward)
        det = oracle.detect(im.astype(np.float64), sigma)
        x = edges[off[b]:off[b + 1]]
        # match GPU rows to oracle rows by position (the sets are identical here)
        eidx = det["eidx"]
        assert len(eidx) == n and np.abs(det["edges"][:, :2] - x[:, :2]).max() < 1e-3
        L, oc = oracle.frozen_loss(im.astype(np.float64), sigma, eidx, w)
        g_ref, s_ref = oracle.frozen_grad(im.astype(np.float64), sigma, eidx, w, oc)
        gs_ref += s_ref
        scale = np.abs(g_ref).max()
        err = np.abs(g[b] - g_ref).max() / scale
        print("sigma %g image %d: %d edgels, max |g - g_oracle| / max|g| = %.2e" % (sigma, b, n, err))
        assert err < 1e-4
        # central finite differences on 64 pixels, step 1e-2 grey levels
        cand = np.flatnonzero(np.abs(g_ref) > 1e-2 * scale)
        pix = np.random.default_rng(b).choice(cand, size=min(64, len(cand)), replace=False)
        base = im.astype(np.float64)
        h = 1e-2
        for k in pix:
            i, j = divmod(int(k), im.shape[1])
            p, m = base.copy(), base.copy()
            p[i, j] += h
            m[i, j] -= h
            fd = (oracle.frozen_loss(p, sigma, eidx, w, octants=oc)[0] -
                  oracle.frozen_loss(m, sigma, eidx, w, octants=oc)[0]) / (2 * h)
            assert abs(fd - g[b][i, j]) <= 1e-4 * scale + 1e-4 * abs(fd), (i, j, fd, g[b][i, j])
    # sigma: FD with the window radii frozen, step 1e-4
    R0, R1 = oracle.radius(sigma)
    fd_s = 0.0
    for b, im in enumerate(imgs):
        det = oracle.detect(im.astype(np.float64), sigma)
        L, oc = oracle.frozen_loss(im.astype(np.float64), sigma, det["eidx"], w)
        fd_s += (oracle.frozen_loss(im.astype(np.float64), sigma + 1e-4, det["eidx"], w, R0, R1, octants=oc)[0] -
                 oracle.frozen_loss(im.astype(np.float64), sigma - 1e-4, det["eidx"], w, R0, R1, octants=oc)[0]) / 2e-4
    print("sigma %g: dL/dsigma gpu %.6f oracle %.6f fd %.6f" % (sigma, gs, gs_ref, fd_s))
    assert gs == pytest.approx(gs_ref, rel=1e-4)
    assert gs == pytest.approx(fd_s, rel=1e-4)


def test_grad_linear_in_upstream(pkg, detector, scene):
    imgs = [scene.shapes_scene(70, 80, 100)]
    w = np.array([0.3, -1.2, 0.7, 2.0])
    _, _, g1, s1 = _forward_backward(pkg, imgs, 1.5, w)
    _, _, g2, s2 = _forward_backward(pkg, imgs, 1.5, 3 * w)
    np.testing.assert_allclose(g2, 3 * g1, rtol=1e-4, atol=1e-6 * np.abs(g1).max())
    assert s2 == pytest.approx(3 * s1, rel=1e-5)


def test_grad_zero_without_edgels(pkg, detector):
    imgs = [np.full((64, 64), 100, np.uint8)]
    e, off, g, gs = _forward_backward(pkg, imgs, 2.0, np.ones(4))
    assert len(e) == 0 and np.all(g == 0) and gs == 0.0


def test_tile_and_index_backward_agree(pkg, detector, scene):
    """The tile-driven backward (plan staging, LDS patches) and the
    index-driven one (used when the plan has run again since the forward)
    give the same gradients."""
    x = torch.from_numpy(np.stack([scene.shapes_scene(80 + k, 120, 150) for k in range(3)]).astype(np.float32))
    x = x.cuda().requires_grad_(True)
    sig = torch.tensor(2.0, dtype=torch.float64, requires_grad=True)
    w = torch.tensor([0.4, -0.8, 1.1, 0.6], device="cuda")
    e1, _ = pkg.detect_differentiable(x, sig)
    (e1 * w).sum().backward()
    g_tile, s_tile = x.grad.clone(), float(sig.grad)
    x.grad = None
    sig.grad = None
    e2, _ = pkg.detect_differentiable(x, sig)
    e3, _ = pkg.detect_differentiable(x, sig)     # the plan runs again: e2's backward takes the index path
    (e2 * w).sum().backward()
    g_idx, s_idx = x.grad.clone(), float(sig.grad)
    assert torch.equal(e1, e2)
    scale = g_tile.abs().max().item()
    assert (g_tile - g_idx).abs().max().item() <= 1e-5 * scale
    assert s_tile == pytest.approx(s_idx, rel=1e-9)


def test_full_batch_64x512_tile_and_index_agree(pkg, detector, scene):
    """The differentiable workload at its full size (64 scenes of 512 x 512
    f32, sigma 1.5, loss weights fixed from a seed): the tile-driven and the
    index-driven backward give the same gradients, and the loss gradient is
    nonzero on every image."""
    x = torch.from_numpy(np.stack([scene.shapes_scene(9000 + b, 512, 512) for b in range(64)]).astype(np.float32))
    x = x.cuda().requires_grad_(True)
    sig = torch.tensor(1.5, dtype=torch.float64, requires_grad=True)
    w = torch.tensor(scene.loss_weights(12345), dtype=torch.float32, device="cuda")
    e1, _ = pkg.detect_differentiable(x, sig)
    (e1 * w).sum().backward()
    g_tile, s_tile = x.grad.clone(), float(sig.grad)
    x.grad = None
    sig.grad = None
    e2, _ = pkg.detect_differentiable(x, sig)
    e3, _ = pkg.detect_differentiable(x, sig)
    (e2 * w).sum().backward()
    g_idx, s_idx = x.grad.clone(), float(sig.grad)
    assert torch.equal(e1, e2) and e1.shape[0] > 500000
    scale = g_tile.abs().max().item()
    assert (g_tile - g_idx).abs().max().item() <= 1e-5 * scale
    # FP64 sums of the same 550k per-edgel terms in two orders
    assert s_tile == pytest.approx(s_idx, rel=1e-7)
    assert bool((g_tile.abs().amax(dim=(1, 2)) > 0).all())


def test_full_batch_64x512_against_oracle(pkg, detector, scene, oracle):
    """The differentiable workload at its full size (64 scenes of 512 x 512 f32,
    sigma 1.5, the bench's loss weights): images 0 and 37 of the batch have the
    oracle's edgel set, their dL/dimage equals the FP64 oracle gradient of the
    frozen loss (an independent route: local central differences and transposed
    filters) to 1e-4 of its largest entry, and central differences of the frozen
    loss agree on 16 pixels of image 0."""
    imgs = [scene.shapes_scene(9000 + b, 512, 512) for b in range(64)]
    w = scene.loss_weights(12345)
    print("64 x 512 x 512 forward + backward", flush=True)
    edges, off, g, _ = _forward_backward(pkg, imgs, 1.5, w)
    assert off[-1] > 500000
    for b in (0, 37):
        im = imgs[b].astype(np.float64)
        det = oracle.detect(im, 1.5)
        x = edges[off[b]:off[b + 1]]
        eidx = det["eidx"]
        assert len(eidx) == len(x) > 5000 and np.abs(det["edges"][:, :2] - x[:, :2]).max() < 1e-3
        L, oc = oracle.frozen_loss(im, 1.5, eidx, w)
        g_ref, _ = oracle.frozen_grad(im, 1.5, eidx, w, oc)
        scale = np.abs(g_ref).max()
        err = np.abs(g[b] - g_ref).max() / scale
        print("image %d: %d edgels, max |g - g_oracle| / max|g| = %.2e" % (b, len(eidx), err), flush=True)
        assert err < 1e-4
        if b == 0:
            cand = np.flatnonzero(np.abs(g_ref) > 1e-2 * scale)
            h = 1e-2
            for k in np.random.default_rng(3).choice(cand, size=16, replace=False):
                i, j = divmod(int(k), im.shape[1])
                p, m = im.copy(), im.copy()
                p[i, j] += h
                m[i, j] -= h
                fd = (oracle.frozen_loss(p, 1.5, eidx, w, octants=oc)[0] -
                      oracle.frozen_loss(m, 1.5, eidx, w, octants=oc)[0]) / (2 * h)
                assert abs(fd - g[b][i, j]) <= 1e-4 * scale + 1e-4 * abs(fd), (i, j, fd, g[b][i, j])


def test_new_tensor_each_step_reuses_the_plan(pkg, detector, scene):
    """A training loop passes a new input tensor every step: the cached plan
    (keyed by shape, strides and parameters, not by address) is re-pointed at
    it instead of rebuilt, at most two plans are kept, and each step's rows
    and gradients equal those of a call on a fresh copy of the same data."""
    autograd = __import__(pkg.__name__ + ".autograd", fromlist=["_PLANS"])
    base = np.stack([scene.shapes_scene(60 + k, 96, 128) for k in range(2)]).astype(np.float32)
    w = torch.tensor([0.3, -0.7, 1.3, 0.5], device="cuda")
    sig = torch.tensor(1.5, dtype=torch.float64)
    plans = set()
    ref_e, ref_g = None, None
    for step in range(4):
        x = torch.from_numpy(base + (step % 2) * 1.0).cuda().requires_grad_(True)   # a new tensor each step
        e, _ = pkg.detect_differentiable(x, sig, detector=detector)
        (e * w).sum().backward()
        plans.add(id(next(reversed(autograd._PLANS.values()))))
        if step % 2 == 0:
            if ref_e is None:
                ref_e, ref_g = e.detach().clone(), x.grad.clone()
            else:
                # rows bit for bit; gradients up to the order of the float atomics that sum them
                assert torch.equal(e, ref_e)
                assert (x.grad - ref_g).abs().max().item() <= 1e-5 * ref_g.abs().max().item()
    assert len(plans) == 1 and len(autograd._PLANS) <= 2
    # a call on another shape adds one plan; the cache never holds more than two
    y = torch.from_numpy(base[:, :80, :100].copy()).cuda()
    pkg.detect_differentiable(y, sig, detector=detector)
    pkg.detect_differentiable(y[:1].clone(), sig, detector=detector)
    assert len(autograd._PLANS) <= 2


def test_in_place_edit_before_backward_is_rejected(pkg, detector, scene):
    """The input is saved for the backward: editing it in place between the
    forward and the backward is an error (torch's version check), not a
    silent gradient of the edited image."""
    x = torch.from_numpy(np.stack([scene.shapes_scene(70, 96, 128)]).astype(np.float32)).cuda()
    xr = x.clone().requires_grad_(True)
    y = xr * 1.0                       # a non-leaf input, edited in place below
    e, _ = pkg.detect_differentiable(y, 2.0, detector=detector)
    with torch.no_grad():
        y.add_(1.0)
    with pytest.raises(RuntimeError):
        e.sum().backward()


def test_device_sigma_gradient_and_updates(pkg, detector, scene):
    """sigma held in device memory: its gradient equals the host sigma's and
    stays on the device, and an in-place update of it between calls is seen
    by the next forward (the wrapper reads it anew when its version changes)."""
    x = torch.from_numpy(np.stack([scene.shapes_scene(90, 96, 128)]).astype(np.float32)).cuda().requires_grad_(True)
    w = torch.tensor([0.2, -0.6, 0.9, 0.4], device="cuda")
    sig_h = torch.tensor(1.5, dtype=torch.float64, requires_grad=True)
    e_h, _ = pkg.detect_differentiable(x, sig_h, detector=detector)
    (e_h * w).sum().backward()
    g_h, s_h = x.grad.clone(), float(sig_h.grad)
    x.grad = None
    sig_d = torch.tensor(1.5, dtype=torch.float64, device="cuda", requires_grad=True)
    e_d, _ = pkg.detect_differentiable(x, sig_d, detector=detector)
    (e_d * w).sum().backward()
    assert sig_d.grad.is_cuda
    assert torch.equal(e_d, e_h)
    assert (x.grad - g_h).abs().max().item() <= 1e-5 * g_h.abs().max().item()
    assert float(sig_d.grad) == pytest.approx(s_h, rel=1e-9)
    with torch.no_grad():
        sig_d.fill_(2.0)
        e2, _ = pkg.detect_differentiable(x, sig_d, detector=detector)
        e2_ref, _ = pkg.detect_differentiable(x, 2.0, detector=detector)
    assert torch.equal(e2, e2_ref) and not torch.equal(e2[:5], e_d[:5])
