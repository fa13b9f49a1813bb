"""CPU tests of the FP64 oracle of the subpixel third-order correction path
(oracle/subpix_oracle.c, restating subpix_TO_correction/subpix_to_correction.cpp).

Pins: the golden vectors of the reference's own subpix_to_correction.cpp
(tests/golden/subpix_ref_*.npz, written by tests/golden/make_golden.py), and
independent mathematics: circular convolution = FFT convolution, Keys cubic
(a = -1/2) reproduces quadratics, the Dollar orientation estimate of a straight
ridge has the closed form mod(atan(tan(phi)|tan(phi)|), pi), the parabola
fit is exact on a quadratic ridge, the third-order orientation of a straight
ridge is its tangent atan2(-cos(phi), sin(phi)), and the isolated-edgel rule
(radius, <= comparison, fewer than two edgels kept).
"""
import glob
import os

import numpy as np
import pytest

import oracle

GOLDEN = sorted(glob.glob(os.path.join(os.path.dirname(__file__), "golden", "subpix_ref_*.npz")))
# the maps of the class constructed with n = 0 (the differentiable form covers those)
GOLDEN0 = [g for g in GOLDEN if "interp_n" not in np.load(g).files]


def test_golden_files_present():
    assert len(GOLDEN) >= 5


@pytest.mark.parametrize("path", GOLDEN, ids=[os.path.basename(p)[11:-4] for p in GOLDEN])
def test_oracle_matches_reference_golden(path):
    z = np.load(path)
    thr, sig, nbr = z["params"]
    theta = z["theta"] if "theta" in z.files else None
    interp_n = int(z["interp_n"]) if "interp_n" in z.files else 0
    o = oracle.subpix(z["E"].astype(np.float64), theta, thresh=thr, sigma=sig, nbr_radius=nbr, nthreads=4,
                      maps=True, interp_n=interp_n)
    assert o["n"] == len(z["edges"]) > 100
    assert o["TO_E"].shape == z["E"].shape
    # same arithmetic in the same order: bit-identical lists and maps
    np.testing.assert_array_equal(o["edges"], z["edges"])
    np.testing.assert_array_equal(o["TO_E"], z["TO_E"])
    np.testing.assert_array_equal(o["TO_O"], z["TO_O"])
    # raster order of the map pixels
    assert np.all(np.diff(o["idx"]) > 0)


def test_interp_goldens_present():
    assert sum("interp_n" in np.load(g).files for g in GOLDEN) >= 3


def test_resize_reproduces_bilinear_functions():
    """The work grid of interp_n = n samples the map at (i / 2^n, j / 2^n):
    corners aligned, original pixels kept exactly, and any function
    a + b y + c x + d x y reproduced between them."""
    rng = np.random.default_rng(11)
    h, w = 7, 9
    a, b, c, d = rng.standard_normal(4)
    yy, xx = np.mgrid[0:h, 0:w].astype(np.float64)
    src = a + b * yy + c * xx + d * xx * yy
    for n in (1, 2, 3):
        dh, dw = oracle.work_shape(h, w, n)
        assert (dh, dw) == ((h - 1) * 2 ** n + 1, (w - 1) * 2 ** n + 1)
        out = oracle.resize(src, dh, dw)
        s = 2 ** n
        np.testing.assert_array_equal(out[::s, ::s], src)
        Y, X = np.mgrid[0:dh, 0:dw].astype(np.float64) / s
        np.testing.assert_allclose(out, a + b * Y + c * X + d * X * Y, rtol=0, atol=1e-12)
        # float rounding of the work map, as the GPU stores it
        np.testing.assert_array_equal(oracle.resize(src, dh, dw, f32=True), out.astype(np.float32))
    # a single row / column stays put
    np.testing.assert_array_equal(oracle.resize(src[:1], 1, 17)[0, ::2], src[0])


def test_interp_zero_is_the_plain_path():
    z = np.load(GOLDEN[0])
    E = z["E"].astype(np.float64)
    a = oracle.subpix(E, None, 0.1, 2.0, 2.0, maps=True)
    b = oracle.subpix(E, None, 0.1, 2.0, 2.0, maps=True, interp_n=0, work_f32=True)
    np.testing.assert_array_equal(a["edges"], b["edges"])    # the map is float already
    np.testing.assert_array_equal(a["TO_E"], b["TO_E"])


def test_interp_straight_ridge_keeps_its_place():
    """A straight vertical ridge at column c of the map: on the work grid of
    interp_n = 1 the edgels lie near x = c in map coordinates (the list is
    divided by 2^n; the parabola on the piecewise-linear upsampled profile is
    biased by about 0.15 px here), on every half row, with the ridge's tangent
    orientation, and each map row's TO pixel holds one of them."""
    h, w, c = 48, 40, 19.25
    xx = np.arange(w, dtype=np.float64)
    E = np.tile(np.exp(-0.5 * ((xx - c) / 1.2) ** 2), (h, 1))
    for n in (0, 1):
        o = oracle.subpix(E, None, 0.1, 2.0, 2.0, maps=True, interp_n=n)
        e = o["edges"]
        assert o["n"] > 20
        np.testing.assert_allclose(e[:, 0], c, atol=0.2)
        np.testing.assert_array_equal(np.diff(e[:, 1]), 0.5 ** n)
        np.testing.assert_allclose(np.abs(np.sin(e[:, 2])), 1.0, atol=1e-6)    # tangent along y: +-pi/2
        rows = np.unique(np.round(e[:, 1]).astype(int))
        assert np.all(o["TO_E"][rows].astype(bool).sum(1) == 1)


def test_taps_closed_forms():
    r, G, dG, d2G = oracle.subpix_taps(2.0)
    assert r == 8 and len(G) == 17
    k = np.arange(-8, 9, dtype=np.float64)
    g = np.exp(-k ** 2 / 8.0) / (np.sqrt(2 * np.pi) * 2.0)
    np.testing.assert_allclose(G, g, rtol=1e-15)
    np.testing.assert_allclose(dG, -k / 4.0 * g, rtol=1e-14, atol=1e-18)
    np.testing.assert_allclose(d2G, (k ** 2 - 4.0) / 16.0 * g, rtol=1e-14, atol=1e-18)
    assert oracle.subpix_taps(1.5)[0] == 6 and oracle.subpix_taps(1.0)[0] == 4


def test_circular_convolution_is_fft_convolution():
    rng = np.random.default_rng(3)
    h, w, r = 24, 30, 3
    src = rng.standard_normal((h, w))
    kx, ky = rng.standard_normal(2 * r + 1), rng.standard_normal(2 * r + 1)
    K = np.zeros((h, w))
    for p in range(-r, r + 1):
        for q in range(-r, r + 1):
            K[p % h, q % w] = ky[p + r] * kx[q + r]
    ref = np.real(np.fft.ifft2(np.fft.fft2(src) * np.fft.fft2(K)))
    np.testing.assert_allclose(oracle.subpix_conv_circ(src, kx, ky), ref, atol=1e-12)
    # a transposed operand fails: the x kernel and the y kernel are not interchangeable
    assert np.abs(oracle.subpix_conv_circ(src, ky, kx) - ref).max() > 1e-3


def test_keys_cubic_reproduces_quadratics():
    h, w = 20, 22
    yy, xx = np.mgrid[0:h, 0:w].astype(np.float64)
    f = lambda x, y: 0.3 + 1.2 * x - 0.7 * y + 0.05 * x * x - 0.02 * x * y + 0.04 * y * y  # noqa: E731
    src = f(xx, yy)
    rng = np.random.default_rng(1)
    for x, y in rng.uniform(2.0, 16.0, (50, 2)):
        assert abs(oracle.subpix_interp_cubic(src, x, y) - f(x, y)) < 1e-11
    # interpolating: node values are returned exactly; cubics are not reproduced
    assert oracle.subpix_interp_cubic(src, 5.0, 7.0) == src[7, 5]
    cub = xx ** 3
    assert abs(oracle.subpix_interp_cubic(cub, 7.3, 5.0) - 7.3 ** 3) > 1e-3
    # clamped borders: sampling beyond the last column repeats it
    assert abs(oracle.subpix_interp_cubic(src, w - 1.0, 3.0) - src[3, w - 1]) < 1e-12


def _ridge(phi, h, w, s, c=(31.7, 32.2)):
    yy, xx = np.mgrid[0:h, 0:w].astype(np.float64)
    u = (xx - c[0]) * np.cos(phi) + (yy - c[1]) * np.sin(phi)
    return np.exp(-u ** 2 / (2 * s * s)), u, xx, yy


@pytest.mark.parametrize("deg", [0, 20, 45, 70, 110, 160])
def test_estimated_orientation_closed_form(deg):
    # for E = g(n . p) the smoothed Hessian is g'' n n^T: Oyy/Oxx = tan^2(phi),
    # sign(-Oxy) = sign(tan(phi)) on the ridge line (g'' < 0)
    phi = np.deg2rad(deg)
    E, u, xx, yy = _ridge(phi, 64, 64, 8.0)
    th = oracle.subpix_estimate_orientation(E)
    sel = (np.abs(u) < 0.5) & (xx > 14) & (xx < 50) & (yy > 14) & (yy < 50)
    t = np.tan(phi) * abs(np.tan(phi))
    want = np.mod(np.arctan(t), np.pi)
    err = np.abs(np.angle(np.exp(2j * (th[sel] - want)))) / 2
    assert sel.sum() > 20 and err.max() < 2.5e-3
    assert np.all((th >= 0) & (th < np.pi))


def test_parabola_exact_on_quadratic_ridge():
    h, w, x0 = 40, 40, 20.3
    yy, xx = np.mgrid[0:h, 0:w].astype(np.float64)
    E = np.maximum(0.0, 0.9 - 0.02 * (xx - x0) ** 2)
    for theta in (0.0, np.pi, -np.pi, 2 * np.pi):
        o = oracle.subpix(E, np.full((h, w), theta), thresh=0.2, sigma=2.0, nbr_radius=2.0)
        e = o["edges"]
        assert o["n"] == h - 9                       # rows [4, h - 5), one edgel each
        np.testing.assert_allclose(e[:, 0], x0, atol=1e-12)
        np.testing.assert_allclose(e[:, 1], np.arange(4, h - 5), atol=0)
        np.testing.assert_allclose(e[:, 3], 0.9, atol=1e-12)
        # tangent of a vertical ridge, with F = -(1, 0) at the crest: -pi/2
        np.testing.assert_allclose(e[:, 2], -np.pi / 2, atol=1e-9)


@pytest.mark.parametrize("deg", [0, 20, 45, 70, 110, 160])
def test_third_order_orientation_of_straight_ridge(deg):
    # F = (g~'^2 + g g~'') n for E = g(n . p): the orientation is the tangent,
    # with the sign of the crest term (negative): atan2(-cos(phi), sin(phi))
    phi = np.deg2rad(deg)
    E, u, xx, yy = _ridge(phi, 64, 64, 2.0)
    E *= 0.9
    o = oracle.subpix(E, np.full(E.shape, phi), thresh=0.2, sigma=2.0, nbr_radius=2.0)
    e = o["edges"]
    sel = (e[:, 0] > 16) & (e[:, 0] < 48) & (e[:, 1] > 16) & (e[:, 1] < 48)
    assert sel.sum() > 20
    want = np.arctan2(-np.cos(phi), np.sin(phi))
    assert np.abs(np.angle(np.exp(1j * (e[sel, 2] - want)))).max() < 2e-4
    uu = (e[sel, 0] - 31.7) * np.cos(phi) + (e[sel, 1] - 32.2) * np.sin(phi)
    assert np.abs(uu).max() < 0.03


def _blobs(centres, h=48, w=48, s=0.6):
    yy, xx = np.mgrid[0:h, 0:w].astype(np.float64)
    E = np.zeros((h, w))
    for cx, cy in centres:
        E = np.maximum(E, np.exp(-((xx - cx) ** 2 + (yy - cy) ** 2) / (2 * s * s)))
    return E


def test_isolated_edgel_rule():
    th = lambda E: np.zeros_like(E)  # noqa: E731  (direction (1, 0))
    # one blob -> one edgel (its centre row), kept: fewer than two edgels
    E1 = _blobs([(20.0, 20.0)])
    o1 = oracle.subpix(E1, th(E1), thresh=0.5, sigma=2.0, nbr_radius=2.0)
    assert o1["n"] == 1
    # two blobs far apart -> both isolated -> none kept; radius 0 keeps both
    E2 = _blobs([(12.0, 12.0), (33.0, 30.0)])
    assert oracle.subpix(E2, th(E2), thresh=0.5, sigma=2.0, nbr_radius=2.0)["n"] == 0
    assert oracle.subpix(E2, th(E2), thresh=0.5, sigma=2.0, nbr_radius=0.0)["n"] == 2
    # a third blob 1.9 px from the second: those two are kept, the first dropped
    E3 = _blobs([(12.0, 12.0), (33.0, 30.0), (33.0 + 1.9, 30.0)], s=0.45)
    o3 = oracle.subpix(E3, th(E3), thresh=0.5, sigma=2.0, nbr_radius=2.0)
    assert o3["n"] == 2 and np.all(o3["edges"][:, 1] == 30.0)
    d = o3["edges"][1, 0] - o3["edges"][0, 0]
    # the same pair is dropped once the radius is below their distance
    assert oracle.subpix(E3, th(E3), thresh=0.5, sigma=2.0, nbr_radius=d * 0.999)["n"] == 0
    assert oracle.subpix(E3, th(E3), thresh=0.5, sigma=2.0, nbr_radius=d)["n"] == 2


def test_maps_last_writer_in_list_order():
    z = np.load([g for g in GOLDEN if "512x512" in g][0])
    thr, sig, nbr = z["params"]
    o = oracle.subpix(z["E"].astype(np.float64), z["theta"] if "theta" in z.files else None, thresh=thr,
                      sigma=sig, nbr_radius=nbr, maps=True)
    e = o["edges"]
    X = np.clip(np.round(e[:, 0]).astype(int), 0, z["E"].shape[1] - 1)
    Y = np.clip(np.round(e[:, 1]).astype(int), 0, z["E"].shape[0] - 1)
    TE = np.zeros_like(o["TO_E"])
    for k in range(len(e)):                       # later rows overwrite earlier ones
        TE[Y[k], X[k]] = e[k, 3]
    np.testing.assert_array_equal(TE, o["TO_E"])
    assert (o["TO_E"] != 0).sum() == len(e) - 1    # one collision in this map


@pytest.mark.parametrize("path", GOLDEN0, ids=[os.path.basename(p)[11:-4] for p in GOLDEN0])
def test_frozen_loss_reproduces_reference_at_forward_point(path):
    """oracle.SubpixFrozen, the finite-difference reference of the GPU adjoint,
    is the forward path with its discrete decisions held: at the forward point
    its outputs are the reference's own list (subpix_to_correction.cpp:623-736,
    the golden rows) to 1e-12, its loss is sum w . rows, and moving the map
    moves the rows (they are recomputed from E, not replayed)."""
    z = np.load(path)
    thr, sig, nbr = z["params"]
    theta = z["theta"] if "theta" in z.files else None
    E = z["E"].astype(np.float64)
    ref = oracle.subpix(E, theta, thresh=thr, sigma=sig, nbr_radius=nbr)
    np.testing.assert_array_equal(ref["edges"], z["edges"])
    fr = oracle.SubpixFrozen(E, ref["idx"], sig, theta)
    w = np.random.default_rng(1).standard_normal((ref["n"], 4))
    L, outs = fr.loss(E, w, outputs=True)
    assert outs.shape == z["edges"].shape
    np.testing.assert_allclose(outs, z["edges"], rtol=0, atol=1e-12)
    assert abs(L - float((w * outs).sum())) <= 1e-12 * float(np.abs(w * outs).sum())
    # one shared weight row gives the same loss as that row repeated
    w1 = np.array([0.3, -1.1, 0.8, 2.0])
    assert abs(fr.loss(E, w1) - fr.loss(E, np.tile(w1, (ref["n"], 1)))) <= 1e-12 * abs(fr.loss(E, w1))
    # a perturbed map gives other rows, of the size of the perturbation
    d = np.random.default_rng(2).standard_normal(E.shape)
    _, outs2 = fr.loss(E + 1e-6 * d, w, outputs=True)
    delta = np.abs(outs2 - outs).max()
    assert 1e-10 < delta < 1e-2


def test_frozen_loss_given_orientation_is_constant_direction():
    """With a given orientation map the direction is an input: the frozen
    rows do not follow the map's Hessian (the estimated-orientation form on the
    same map gives other rows)."""
    z = np.load([g for g in GOLDEN if "orient" in g][0])
    thr, sig, nbr = z["params"]
    E = z["E"].astype(np.float64)
    ref = oracle.subpix(E, z["theta"], thresh=thr, sigma=sig, nbr_radius=nbr)
    given = oracle.SubpixFrozen(E, ref["idx"], sig, z["theta"])
    est = oracle.SubpixFrozen(E, ref["idx"], sig, None)
    w = np.ones(4)
    _, a = given.loss(E, w, outputs=True)
    _, b = est.loss(E, w, outputs=True)
    np.testing.assert_allclose(a, z["edges"], rtol=0, atol=1e-12)
    assert np.abs(a - b).max() > 1e-3
