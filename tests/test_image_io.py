"""Image decoding of the command-line drop-in (toed_image_load, host code, no
GPU): PGM (P5) and PNG written here from known pixels, including every PNG
row filter, colour and 16-bit layouts; the reference's no-OpenCV PGM read
(one byte early); and, where the reference tree is present, its own sample
images against the independently decoded pixels of the golden fixtures."""
import os
import struct
import zlib

import numpy as np
import pytest

GOLD = os.path.join(os.path.dirname(__file__), "golden")
REF_IMAGES = "/root/reference/input_images"


def _png(path, arr, color_type, depth=8, filters=(0, 1, 2, 3, 4)):
    """Minimal PNG encoder: arr is H x W x channels (uint8 or uint16); row y
    uses filter filters[y % len(filters)]."""
    arr = np.asarray(arr)
    if arr.ndim == 2:
        arr = arr[:, :, None]
    H, W, ch = arr.shape
    if depth == 16:
        raw = arr.astype(">u2").view(np.uint8).reshape(H, W * ch * 2).astype(np.int32)
    else:
        raw = arr.astype(np.uint8).reshape(H, W * ch).astype(np.int32)
    bpp = ch * depth // 8
    out = bytearray()
    prev = np.zeros(raw.shape[1], np.int32)
    for y in range(H):
        f = filters[y % len(filters)]
        cur = raw[y]
        a = np.concatenate([np.zeros(bpp, np.int32), cur[:-bpp]])
        c = np.concatenate([np.zeros(bpp, np.int32), prev[:-bpp]])
        b = prev
        if f == 0:
            pred = np.zeros_like(cur)
        elif f == 1:
            pred = a
        elif f == 2:
            pred = b
        elif f == 3:
            pred = (a + b) // 2
        else:
            p = a + b - c
            pa, pb, pc = np.abs(p - a), np.abs(p - b), np.abs(p - c)
            pred = np.where((pa <= pb) & (pa <= pc), a, np.where(pb <= pc, b, c))
        out.append(f)
        out += ((cur - pred) & 255).astype(np.uint8).tobytes()
        prev = cur

    def chunk(t, d):
        return struct.pack(">I", len(d)) + t + d + struct.pack(">I", zlib.crc32(t + d) & 0xFFFFFFFF)

    data = b"\x89PNG\r\n\x1a\n" + chunk(b"IHDR", struct.pack(">IIBBBBB", W, H, depth, color_type, 0, 0, 0))
    data += chunk(b"IDAT", zlib.compress(bytes(out), 6)) + chunk(b"IEND", b"")
    open(path, "wb").write(data)


def _pgm(path, arr, header=b"P5\n# a comment line\n%d %d\n255\n"):
    H, W = arr.shape
    open(path, "wb").write((header % (W, H)) + arr.astype(np.uint8).tobytes())


@pytest.fixture(scope="module")
def pixels():
    rng = np.random.default_rng(17)
    return rng.integers(0, 256, size=(37, 53), dtype=np.uint8)


def test_pgm_p5(pkg, tmp_path, pixels):
    p = str(tmp_path / "a.pgm")
    _pgm(p, pixels)
    assert np.array_equal(pkg.load_image(p), pixels)


def test_pgm_reference_raw_offset(pkg, tmp_path, pixels):
    """main_gpu_cpu.cpp:131 + gpu_toed.hpp:192-196: after `>> intensity` the
    stream sits on the newline; get() then yields it first and every pixel
    one position late."""
    p = str(tmp_path / "b.pgm")
    _pgm(p, pixels, header=b"P5\n%d %d\n255\n")
    raw = pkg.load_image(p, pgm_raw_offset=True)
    flat = pixels.reshape(-1)
    assert raw.shape == pixels.shape
    assert raw.reshape(-1)[0] == ord("\n")
    assert np.array_equal(raw.reshape(-1)[1:], flat[:-1])


@pytest.mark.parametrize("filters", [(0,), (1,), (2,), (3,), (4,), (0, 1, 2, 3, 4)])
def test_png_grey_every_filter(pkg, tmp_path, pixels, filters):
    p = str(tmp_path / "g.png")
    _png(p, pixels, 0, filters=filters)
    assert np.array_equal(pkg.load_image(p), pixels)


def test_png_colour_alpha_and_16bit(pkg, tmp_path):
    rng = np.random.default_rng(5)
    rgb = rng.integers(0, 256, size=(21, 30, 3), dtype=np.uint8)
    want = ((rgb[..., 0].astype(np.int64) * 4899 + rgb[..., 1].astype(np.int64) * 9617 +
             rgb[..., 2].astype(np.int64) * 1868 + 8192) >> 14).astype(np.uint8)
    # the 14-bit weights are BT.601's 0.299 / 0.587 / 0.114 to within 2e-5 each: the
    # result is the rounded exact grey level, off by at most 0.5 + 255 * 5e-5
    ref = 0.299 * rgb[..., 0] + 0.587 * rgb[..., 1] + 0.114 * rgb[..., 2]
    assert np.abs(want.astype(np.float64) - ref).max() <= 0.5 + 255 * 5e-5
    p = str(tmp_path / "c.png")
    _png(p, rgb, 2)
    assert np.array_equal(pkg.load_image(p), want)
    rgba = np.concatenate([rgb, rng.integers(0, 256, size=(21, 30, 1), dtype=np.uint8)], axis=2)
    _png(p, rgba, 6)
    assert np.array_equal(pkg.load_image(p), want)
    grey = rgb[..., 0]
    ga = np.stack([grey, rgb[..., 1]], axis=2)
    _png(p, ga, 4)
    assert np.array_equal(pkg.load_image(p), grey)
    g16 = rng.integers(0, 65536, size=(21, 30), dtype=np.uint16)
    _png(p, g16, 0, depth=16)
    assert np.array_equal(pkg.load_image(p), (g16 >> 8).astype(np.uint8))


def test_bad_files_are_errors(pkg, tmp_path):
    p = str(tmp_path / "x.bin")
    open(p, "wb").write(b"not an image at all")
    with pytest.raises(pkg.ToedError):
        pkg.load_image(p)
    with pytest.raises(pkg.ToedError):
        pkg.load_image(str(tmp_path / "missing.pgm"))
    open(p, "wb").write(b"P5\n4 4\n255\n" + bytes(8))   # truncated raster
    with pytest.raises(pkg.ToedError):
        pkg.load_image(p)


@pytest.mark.skipif(not os.path.isdir(REF_IMAGES), reason="reference sample images not present")
@pytest.mark.parametrize("name,fixture", [("2018.pgm", "toed_ref_img_2018_pgm.npz"),
                                          ("t_less_test_img.pgm", "toed_ref_img_t_less_pgm.npz"),
                                          ("euroc_sample_img.png", "toed_ref_img_euroc_png.npz")])
def test_reference_sample_images(pkg, name, fixture):
    """The reference's own sample files decode to the pixels the golden
    fixtures hold (decoded there by an independent Python reader)."""
    img = pkg.load_image(os.path.join(REF_IMAGES, name))
    want = np.load(os.path.join(GOLD, fixture))["img"]
    assert np.array_equal(img, want)
