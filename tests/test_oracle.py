"""The FP64 oracle, pinned against the reference's own output and against
what the mathematics fixes (CPU only)."""
import glob
import math
import os

import numpy as np
import pytest

GOLDEN = sorted(glob.glob(os.path.join(os.path.dirname(__file__), "golden", "toed_ref_*.npz")))


@pytest.mark.parametrize("path", GOLDEN, ids=[os.path.basename(p)[9:-4] for p in GOLDEN])
def test_oracle_matches_reference_golden(oracle, path):
    """Edgel list of the compiled reference cpu_toed.cpp (tests/golden) vs the
    restatement: same edgels in the same order, values to ~1e-12.  The
    reference prints its sigma=2 taps to 15 digits, the restatement evaluates
    the closed forms, hence not bit-exact."""
    d = np.load(path)
    img, ref = d["img"], d["edges"]
    o = oracle.detect(img.astype(np.float64), 2.0, nthreads=8)
    e = o["edges"]
    assert e.shape == ref.shape
    np.testing.assert_allclose(e[:, :2], ref[:, :2], rtol=0, atol=1e-9)
    # orientation: same NaN pattern, angles to 1e-9 (mod the atan2 cut at +-pi)
    assert np.array_equal(np.isnan(e[:, 2]), np.isnan(ref[:, 2]))
    dth = np.abs(np.nan_to_num(e[:, 2] - ref[:, 2]))
    assert np.minimum(dth, np.abs(dth - 2 * np.pi)).max() < 1e-9
    np.testing.assert_allclose(e[:, 3], ref[:, 3], rtol=1e-10, atol=0)


def test_taps_closed_forms(oracle):
    """sigma = 2: R0 = 8 (17-tap window), R1 = 9; sampled Gaussian sums to 1,
    first moment of G' is -1, G'' has zero sum and second moment 2, parity."""
    r0, r1, t = oracle.taps(2.0)
    assert (r0, r1) == (8, 9)
    x = np.arange(-r1, r1 + 1, dtype=np.float64)
    G, G1, G2, G3 = t[0], t[1], t[2], t[3]
    assert abs(G.sum() - 1) < 1e-4
    assert abs((x * G1).sum() + 1) < 1e-3
    assert abs(G2.sum()) < 1e-3 and abs((x * x * G2).sum() - 2) < 1e-2
    assert abs((x ** 3 * G3).sum() + 6) < 5e-2
    np.testing.assert_allclose(G, G[::-1], rtol=0, atol=0)
    np.testing.assert_allclose(G1, -G1[::-1], rtol=0, atol=0)
    np.testing.assert_allclose(G3, -G3[::-1], rtol=0, atol=0)
    # shifted taps sample x + 1/2: symmetric about x = -1/2
    Gs = t[4]
    np.testing.assert_allclose(Gs[:-1], Gs[:-1][::-1], rtol=1e-15)
    # the reference's sigma = 2 literals (cpu_toed.cpp:150-167) at a few points
    assert abs(G[9] - 0.199471140200716) < 1e-15
    assert abs(G1[0] - 1.79817087452687e-05) < 1e-19
    assert abs(t[5][8] - 0.0241667573001781) < 1e-16
    assert abs(t[7][18] + 2.92094529738860e-05) < 1e-19
    for s, rr in [(1.0, (4, 5)), (1.5, (6, 7)), (0.5, (2, 3))]:
        assert oracle.radius(s) == rr


def test_impulse_response(oracle):
    """A unit impulse reproduces the tap products: phase 0 uses the R0 window,
    the other phases R1 with shifted taps (cpu_toed.cpp:205-344)."""
    H = W = 41
    img = np.zeros((H, W))
    ci, cj = 20, 20
    img[ci, cj] = 1.0
    r0, r1, t = oracle.taps(2.0)
    c = oracle.convolve(img, 2.0, resp9=True)
    G, G1 = t[0], t[1]
    Gs, G1s = t[4], t[5]
    for (i, j) in [(20, 20), (23, 18), (12, 28), (20 + 9, 20), (20, 20 - 9), (28, 28)]:
        p, q = i - ci, j - cj   # img(i-p, j-q) = 1
        k, l = p + r1, q + r1
        # phase 0: R0 window -> zero at |p| or |q| == R1
        inside0 = abs(p) <= r0 and abs(q) <= r0
        assert c["Ix"][2 * i, 2 * j] == pytest.approx(G1[l] * G[k] if inside0 else 0.0, abs=1e-18)
        assert c["Iy"][2 * i, 2 * j] == pytest.approx(G[l] * G1[k] if inside0 else 0.0, abs=1e-18)
        # phase (0,1): x shifted
        assert c["Ix"][2 * i, 2 * j + 1] == pytest.approx(G1s[l] * G[k], abs=1e-18)
        # phase (1,0): y shifted
        assert c["Iy"][2 * i + 1, 2 * j] == pytest.approx(G[l] * G1s[k], abs=1e-18)
        # phase (1,1)
        assert c["Ix"][2 * i + 1, 2 * j + 1] == pytest.approx(G1s[l] * Gs[k], abs=1e-18)
        # third-order responses
        assert c["resp9"][7][2 * i, 2 * j + 1] == pytest.approx(t[7][l] * G[k], abs=1e-18)   # fxxx
        assert c["resp9"][8][2 * i + 1, 2 * j] == pytest.approx(G[l] * t[7][k], abs=1e-18)   # fyyy


def test_linear_ramp_and_constant(oracle):
    """I = a*x + b*y + c at an interior pixel, phase 0 (window R0), with the
    window moments S0 = sum G, M1 = -sum q G'(q), S2 = sum G'':
        Ix = a M1 S0,  Iy = b M1 S0,  Ixy = 0,
        Ixx = Iyy = S2 S0 I(i, j),  Ixxy = S2 b M1,  Ixyy = S2 a M1.
    A constant image gives zero gradient."""
    H, W = 60, 70
    a, b = 1.5, -0.75
    yy, xx = np.mgrid[0:H, 0:W].astype(np.float64)
    img = a * xx + b * yy + 20.0
    r0, r1, t = oracle.taps(2.0)
    q = np.arange(-r0, r0 + 1)
    S0 = t[0][q + r1].sum()
    M1 = -(q * t[1][q + r1]).sum()
    S2 = t[2][q + r1].sum()
    c = oracle.convolve(img, 2.0, resp9=True)
    i, j = 30, 35
    o = (2 * i, 2 * j)
    r9 = c["resp9"]
    assert r9[0][o] == pytest.approx(a * M1 * S0, rel=1e-12)
    assert r9[1][o] == pytest.approx(b * M1 * S0, rel=1e-12)
    assert abs(r9[3][o]) < 1e-12
    assert r9[2][o] == pytest.approx(S2 * S0 * img[i, j], rel=1e-9)
    assert r9[4][o] == pytest.approx(S2 * S0 * img[i, j], rel=1e-9)
    assert r9[5][o] == pytest.approx(S2 * b * M1, rel=1e-9)
    assert r9[6][o] == pytest.approx(S2 * a * M1, rel=1e-9)
    c0 = oracle.convolve(np.full((H, W), 77.0), 2.0)
    assert abs(c0["Ix"][2 * i, 2 * j]) < 1e-12 and abs(c0["Iy"][2 * i, 2 * j]) < 1e-12


def test_transpose_and_translation(oracle, scene):
    """Transposing the image swaps x and y (Ix <-> Iy on the transposed interp
    grid); an integer shift away from the borders shifts every plane."""
    img = scene.shapes_scene(11, 60, 76).astype(np.float64)
    c = oracle.convolve(img, 2.0)
    ct = oracle.convolve(np.ascontiguousarray(img.T), 2.0)
    np.testing.assert_allclose(ct["Ix"], c["Iy"].T, rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(ct["Iy"], c["Ix"].T, rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(ct["mag"], c["mag"].T, rtol=1e-12, atol=1e-12)
    big = np.zeros((100, 110))
    big[20:80, 17:93] = img
    cb = oracle.convolve(big, 2.0)
    inner = (slice(2 * 10, 2 * 50), slice(2 * 10, 2 * 60))   # >= R1 from the pasted image border
    sh = (slice(2 * 30, 2 * 70), slice(2 * 27, 2 * 77))
    np.testing.assert_allclose(cb["Ix"][sh], c["Ix"][inner], rtol=1e-12, atol=1e-10)
    np.testing.assert_allclose(cb["mag"][sh], c["mag"][inner], rtol=1e-12, atol=1e-10)


def _step_image(H, W, u, lo=50.0, hi=180.0):
    """Vertical step at continuous x = u (pixel j covers [j, j+1))."""
    x = np.arange(W, dtype=np.float64)
    cover = np.clip(x + 1 - u, 0, 1)   # fraction of pixel j right of u
    row = lo + (hi - lo) * cover
    return np.tile(row, (H, 1))


@pytest.mark.parametrize("u", [40.0, 40.3, 40.5, 40.77])
def test_step_edge_subpixel(oracle, u):
    """A straight step at x = u yields one edgel column per interp row at
    x_out = u - 1 (pixel centres at j + 1/2 and the (x_interp - 1)/2 output
    convention, cpu_toed.cpp:516) within the parabola fit's bias, with the
    orientation along the edge (+-pi/2) and strength close to the peak."""
    img = _step_image(64, 80, u)
    o = oracle.detect(img, 2.0)
    e = o["edges"]
    rows = o["eidx"] // 160
    sel = (rows >= 20) & (rows < 108)
    e = e[sel]
    assert len(e) == int(sel.sum()) > 50
    assert np.abs(e[:, 0] - (u - 1)).max() < 0.03
    # the shifted windows span [-R1 + 1/2, R1 + 1/2] (cpu_toed.cpp:164-167), so
    # their G' taps do not cancel exactly: a ~4e-6 rad tilt on odd rows
    assert np.all(np.abs(np.abs(e[:, 2]) - np.pi / 2) < 1e-4)
    _, r1, t = oracle.taps(2.0)
    peak = 130.0 * t[4].max()
    assert np.all(np.abs(e[:, 3] - peak) / peak < 0.1)


def test_disk_orientation_is_tangent(oracle):
    """Edgels on a disk boundary: third-order orientation tangent to the circle,
    theta = phi + pi/2 (mod pi) where phi is the polar angle of the edgel."""
    H = W = 160
    cy, cx, r = 80.3, 79.6, 45.0
    ss = 8
    yy, xx = np.mgrid[0:H * ss, 0:W * ss].astype(np.float64)
    inside = ((xx + 0.5) / ss - cx) ** 2 + ((yy + 0.5) / ss - cy) ** 2 < r * r
    img = 40 + 150 * inside.reshape(H, ss, W, ss).mean(axis=(1, 3))
    o = oracle.detect(img, 2.0)
    e = o["edges"]
    assert len(e) > 300
    # output coordinates have the -1/2 (pixel-centre) and -1/2 (interp) offsets
    px, py = e[:, 0] + 1.0, e[:, 1] + 1.0
    phi = np.arctan2(py - cy, px - cx)
    rad = np.hypot(px - cx, py - cy)
    assert np.abs(rad - r).max() < 0.1
    d = np.mod(e[:, 2] - (phi + np.pi / 2), np.pi)
    d = np.minimum(d, np.pi - d)
    assert np.quantile(d, 0.99) < 0.02


def test_nms_margins_and_border(oracle, scene):
    """The oracle's margin planes are finite exactly where a comparison was
    evaluated and no edgel is reported within the 10-pixel border."""
    img = scene.shapes_scene(5, 70, 90).astype(np.float64)
    o = oracle.detect(img, 2.0, margins=True)
    rows, cols = o["eidx"] // 180, o["eidx"] % 180
    assert rows.min() >= 10 and rows.max() < 140 - 10 and cols.min() >= 10 and cols.max() < 180 - 10
    mm = o["mmag"]
    assert np.isinf(mm[:10]).all() and np.isfinite(mm[10:-10, 10:-10]).all()
    assert np.all(np.diff(o["eidx"]) > 0)


def test_tiny_and_empty_images(oracle):
    # interior [10, 2H-10) is empty for H <= 10: no edgels (reference grid would be empty)
    for (H, W) in [(1, 1), (5, 7), (10, 40)]:
        img = np.random.default_rng(0).uniform(0, 255, (H, W))
        assert oracle.detect(img, 2.0)["n"] == 0


def test_qad_cubic_convolution_reproduces_quadratics(oracle):
    """Keys' kernel with a = -1/2 reproduces polynomials up to degree 2 (and not
    3): the interpolation of the QaD orientation mode."""
    yy, xx = np.mgrid[0:12, 0:12].astype(np.float64)
    quad = 1 + 2 * xx - 0.5 * yy + 0.25 * xx * xx + 0.75 * xx * yy - 0.3 * yy * yy
    maps = np.stack([quad] * 9)
    rng = np.random.default_rng(3)
    X, Y = rng.uniform(3, 8, 50), rng.uniform(3, 8, 50)
    want = 1 + 2 * X - 0.5 * Y + 0.25 * X * X + 0.75 * X * Y - 0.3 * Y * Y
    np.testing.assert_allclose(oracle.qad_responses(maps, X, Y)[4], want, rtol=1e-12, atol=1e-12)
    cub = np.stack([xx ** 3] * 9)
    assert np.abs(oracle.qad_responses(cub, X, Y)[0] - X ** 3).max() > 1e-3


def test_qad_at_the_pixel_is_the_pixel_value(oracle, scene):
    """At integer interp coordinates the interpolation returns the samples, so
    the QaD orientation there is the plain orientation map's value."""
    img = scene.shapes_scene(31, 64, 96).astype(np.float64)
    pl = oracle.convolve(img, 2.0, resp9=True)
    rng = np.random.default_rng(4)
    R, C = rng.integers(20, 108, 40), rng.integers(20, 172, 40)
    acc = oracle.qad_responses(pl["resp9"], C.astype(np.float64), R.astype(np.float64))
    th = np.array([oracle.to_theta(*acc[:, k]) for k in range(len(R))])
    np.testing.assert_allclose(th, pl["orient"][R, C], rtol=0, atol=1e-12)


def test_qad_orientation_of_a_straight_edge(oracle):
    """For a straight edge I = g(n . p) the third-order vector is parallel to n
    everywhere, so the orientation evaluated at the subpixel location is the
    line's direction, as at the pixel, up to the sampled filters' error
    (about 6e-4 rad at sigma 2 in both modes)."""
    for phi in (0.37, 1.1, 2.5):
        yy, xx = np.mgrid[0:96, 0:96].astype(np.float64)
        d = (xx - 48) * np.cos(phi) + (yy - 48) * np.sin(phi)
        img = 40 + 120 / (1 + np.exp(-d / 0.8))
        e = oracle.detect_qad(img, 2.0)["edges"]
        inner = (np.abs(e[:, 0] - 48) < 20) & (np.abs(e[:, 1] - 48) < 20)
        assert inner.sum() > 20
        # orientation is atan2 of the tangent direction: phi + pi/2 modulo pi
        diff = np.angle(np.exp(2j * (e[inner, 2] - (phi + np.pi / 2)))) / 2
        assert np.abs(diff).max() < 1e-3, phi
