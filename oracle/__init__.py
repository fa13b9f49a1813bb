"""TEST INFRASTRUCTURE ONLY: ctypes bindings of the FP64 TOED oracle.

Only tests/, ``__graft_entry__.smoke()`` and bench.py's ``cpu_baseline`` leg
may import this package, and only as the checker (or the timed CPU baseline).
The C sources restate the reference's ``cpu_toed.cpp`` (see the header of
``toed_oracle.c`` for the file:line map) and its
``subpix_TO_correction/subpix_to_correction.cpp`` (``subpix_oracle.c``).
The curvelet formation stage has no restatement here: its checker is the
reference's own header-only curvelet code, compiled by ``make ref`` into
``_ref/`` (``run_curvelet_ref``), and the golden hashes it wrote.
"""
import ctypes
import os
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

_dp = ctypes.POINTER(ctypes.c_double)
_i64p = ctypes.POINTER(ctypes.c_int64)


def build(ref=False):
    """Compile liboracle.so (and, with ref=True, the reference harness)."""
    subprocess.check_call(["make", "-s", "-C", HERE, "all"])
    if ref:
        subprocess.check_call(["make", "-s", "-C", HERE, "ref"])


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        L.toed_oracle_radius.argtypes = [ctypes.c_double, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]
        L.toed_oracle_taps.argtypes = [ctypes.c_double, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int), _dp]
        L.toed_oracle_convolve.argtypes = [_dp, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_int,
                                           _dp, _dp, _dp, _dp, _dp]
        L.toed_oracle_nms.argtypes = [_dp, _dp, _dp, _dp, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                      ctypes.c_int, ctypes.c_int, _dp, ctypes.c_long, _i64p, _dp, _dp]
        L.toed_oracle_nms.restype = ctypes.c_long
        L.toed_oracle_detect.argtypes = [_dp, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_int,
                                         ctypes.c_double, ctypes.c_int, _dp, ctypes.c_long, _i64p,
                                         _dp, _dp, _dp, _dp, _dp, _dp]
        L.toed_oracle_detect.restype = ctypes.c_long
        L.toed_oracle_to_theta.argtypes = [ctypes.c_double] * 9
        L.toed_oracle_to_theta.restype = ctypes.c_double
        _ip = ctypes.POINTER(ctypes.c_int)
        L.toed_oracle_taps_frozen.argtypes = [ctypes.c_double, ctypes.c_int, ctypes.c_int, _dp, _dp]
        L.toed_oracle_frozen_loss.argtypes = [_dp, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_int,
                                              ctypes.c_int, _i64p, ctypes.c_long, _dp, _ip, _dp]
        L.toed_oracle_frozen_loss.restype = ctypes.c_double
        L.toed_oracle_frozen_grad.argtypes = [_dp, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_int,
                                              ctypes.c_int, _i64p, ctypes.c_long, _dp, _ip, _dp, _dp]
        L.toed_subpix_taps.argtypes = [ctypes.c_double, _ip, _dp, _dp, _dp]
        L.toed_subpix_estimate_orientation.argtypes = [_dp, ctypes.c_int, ctypes.c_int, _dp]
        L.toed_subpix_conv_circ.argtypes = [_dp, _dp, ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp, _dp]
        L.toed_subpix_interp_cubic.argtypes = [_dp, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double]
        L.toed_subpix_interp_cubic.restype = ctypes.c_double
        L.toed_subpix_oracle.argtypes = [_dp, ctypes.c_int, ctypes.c_int, _dp, ctypes.c_double, ctypes.c_double,
                                         ctypes.c_double, ctypes.c_int, _dp, ctypes.c_long, _i64p, _dp, _dp]
        L.toed_subpix_oracle.restype = ctypes.c_long
        L.toed_subpix_oracle_n.argtypes = [_dp, ctypes.c_int, ctypes.c_int, _dp, ctypes.c_double, ctypes.c_double,
                                           ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp,
                                           ctypes.c_long, _i64p, _dp, _dp]
        L.toed_subpix_oracle_n.restype = ctypes.c_long
        L.toed_subpix_resize.argtypes = [_dp, ctypes.c_int, ctypes.c_int, _dp, ctypes.c_int, ctypes.c_int,
                                         ctypes.c_int]
        _i8p = ctypes.POINTER(ctypes.c_byte)
        L.toed_subpix_frozen_state.argtypes = [_dp, ctypes.c_int, ctypes.c_int, _dp, _i64p, ctypes.c_long, _i8p,
                                               _i8p, _ip, _ip, _ip, ctypes.c_double]
        L.toed_subpix_frozen_loss.argtypes = [_dp, ctypes.c_int, ctypes.c_int, _dp, ctypes.c_double, _i64p,
                                              ctypes.c_long, _dp, ctypes.c_int, _i8p, _i8p, _ip, _ip, _ip, _dp]
        L.toed_subpix_frozen_loss.restype = ctypes.c_double
        _LIB = L
    return _LIB


def _p(a, t=_dp):
    return None if a is None else a.ctypes.data_as(t)


def radius(sigma):
    r0, r1 = ctypes.c_int(), ctypes.c_int()
    if lib().toed_oracle_radius(float(sigma), ctypes.byref(r0), ctypes.byref(r1)):
        raise ValueError("bad sigma %r" % sigma)
    return r0.value, r1.value


def taps(sigma):
    """(R0, R1, taps[8, 2*R1+1]): G, G', G'', G''' at x=k-R1, then at x+1/2."""
    r0, r1 = radius(sigma)
    t = np.zeros((8, 2 * r1 + 1), np.float64)
    lib().toed_oracle_taps(float(sigma), None, None, _p(t))
    return r0, r1, t


def convolve(img, sigma=2.0, nthreads=0, resp9=False):
    """Ix, Iy, |grad|, theta planes (2H x 2W); optionally the nine responses."""
    img = np.ascontiguousarray(img, np.float64)
    H, W = img.shape
    out = [np.empty((2 * H, 2 * W), np.float64) for _ in range(4)]
    r9 = np.empty((9, 2 * H, 2 * W), np.float64) if resp9 else None
    if lib().toed_oracle_convolve(_p(img), H, W, float(sigma), int(nthreads), *[_p(o) for o in out], _p(r9)):
        raise ValueError("convolve failed")
    res = dict(Ix=out[0], Iy=out[1], mag=out[2], orient=out[3])
    if resp9:
        res["resp9"] = r9
    return res


def nms(planes, thresh=2.0, border=10, nthreads=0, margins=False):
    Ix, Iy, mag, orient = (np.ascontiguousarray(planes[k], np.float64) for k in ("Ix", "Iy", "mag", "orient"))
    IH, IW = Ix.shape
    cap = IH * IW
    edges = np.empty((cap, 4), np.float64)
    eidx = np.empty(cap, np.int64)
    mm = np.empty((IH, IW), np.float64) if margins else None
    ms = np.empty((IH, IW), np.float64) if margins else None
    n = lib().toed_oracle_nms(_p(Ix), _p(Iy), _p(mag), _p(orient), IH // 2, IW // 2, float(thresh), int(border),
                              int(nthreads), _p(edges), cap, _p(eidx, _i64p), _p(mm), _p(ms))
    if n < 0:
        raise MemoryError("oracle nms failed")
    res = dict(edges=edges[:n].copy(), eidx=eidx[:n].copy(), n=int(n))
    if margins:
        res["mmag"], res["msub"] = mm, ms
    return res


def detect(img, sigma=2.0, nthreads=0, thresh=2.0, border=10, planes=False, margins=False):
    """FP64 TOED on a 2-D image (any real dtype).  Returns dict with
    ``edges`` (N x 4: x, y, theta, strength), ``eidx`` (interp index)."""
    img = np.ascontiguousarray(img, np.float64)
    H, W = img.shape
    if planes:
        pl = convolve(img, sigma, nthreads)
        res = nms(pl, thresh, border, nthreads, margins)
        res.update(pl)
        return res
    IH, IW = 2 * H, 2 * W
    cap = IH * IW
    edges = np.empty((cap, 4), np.float64)
    eidx = np.empty(cap, np.int64)
    mm = np.empty((IH, IW), np.float64) if margins else None
    ms = np.empty((IH, IW), np.float64) if margins else None
    n = lib().toed_oracle_detect(_p(img), H, W, float(sigma), int(nthreads), float(thresh), int(border),
                                 _p(edges), cap, _p(eidx, _i64p), None, None, None, None, _p(mm), _p(ms))
    if n < 0:
        raise ValueError("oracle detect failed")
    res = dict(edges=edges[:n].copy(), eidx=eidx[:n].copy(), n=int(n))
    if margins:
        res["mmag"], res["msub"] = mm, ms
    return res


def to_theta(*d):
    return lib().toed_oracle_to_theta(*[float(v) for v in d])


def keys(s):
    """Cubic convolution kernel of Keys with a = -1/2 (MATLAB interp2 'cubic')."""
    x = np.abs(np.asarray(s, np.float64))
    return np.where(x <= 1, (1.5 * x - 2.5) * x * x + 1, np.where(x < 2, ((-0.5 * x + 2.5) * x - 4) * x + 2, 0.0))


def qad_responses(resp9, Xs, Ys):
    """The nine interp-grid response maps (9, IH, IW) interpolated at interp
    coordinates (Xs, Ys) by cubic convolution over the 4 x 4 interp pixels
    around each point (third_order_edge_detector.m:103-112)."""
    Xs, Ys = np.atleast_1d(np.asarray(Xs, np.float64)), np.atleast_1d(np.asarray(Ys, np.float64))
    rx, ry = np.floor(Xs).astype(np.int64), np.floor(Ys).astype(np.int64)
    fx, fy = Xs - rx, Ys - ry
    acc = np.zeros((9, len(Xs)))
    for i in range(4):
        for j in range(4):
            acc += keys(fy + 1 - i) * keys(fx + 1 - j) * resp9[:, ry - 1 + i, rx - 1 + j]
    return acc


def detect_qad(img, sigma=2.0, thresh=2.0, border=10, nthreads=0, margins=False):
    """FP64 detection with the third-order orientation evaluated at each
    edgel's subpixel location (the MATLAB detector's QaD path,
    third_order_edge_detector.m:91-121 and 125-133): the nine response maps of
    cpu_toed.cpp:128-369 interpolated there (qad_responses), then the formula
    of cpu_toed.cpp:227-232.  Positions and strengths are those of detect()."""
    pl = convolve(img, sigma, nthreads, resp9=True)
    res = nms(pl, thresh, border, nthreads, margins)
    e = res["edges"].copy()
    if len(e):
        acc = qad_responses(pl["resp9"], 2 * e[:, 0] + 1, 2 * e[:, 1] + 1)
        e[:, 2] = [to_theta(*acc[:, k]) for k in range(len(e))]
    res["edges"] = e
    res["resp9"] = pl["resp9"]
    res["orient"] = pl["orient"]
    return res


REF_BIN = os.path.join(HERE, "_ref", "toed_ref")


def ref_available():
    return os.path.exists(REF_BIN)


def run_ref(img_u8, nthreads=1, env=None):
    """Run the reference's own cpu_toed.cpp (compiled into oracle/_ref) on a
    uint8 image.  Returns (edges N x 4 float64, time_conv_s, time_nms_s): the
    reference's own omp_get_wtime intervals around its convolution and NMS
    (cpu_toed.cpp:171/359, 390/499).  env: extra environment variables for the
    run (e.g. OMP_PROC_BIND)."""
    img_u8 = np.ascontiguousarray(img_u8, np.uint8)
    H, W = img_u8.shape
    full_env = dict(os.environ, **(env or {}))
    with tempfile.TemporaryDirectory() as td:
        raw = os.path.join(td, "img.raw")
        out = os.path.join(td, "out.bin")
        img_u8.tofile(raw)
        subprocess.run([REF_BIN, raw, str(H), str(W), str(int(nthreads)), out, td],
                       check=True, stdout=subprocess.DEVNULL, env=full_env)
        with open(out, "rb") as f:
            n = int(np.frombuffer(f.read(8), np.int64)[0])
            tc, tn = np.frombuffer(f.read(16), np.float64)
            edges = np.frombuffer(f.read(n * 32), np.float64).reshape(n, 4).copy()
    return edges, float(tc), float(tn)


# ---- differentiable form (toed_grad_oracle.c) ------------------------------

def taps_frozen(sigma, R0, R1):
    """(taps[8, 2R1+1], d taps / d sigma) at sigma with frozen window radii."""
    n = 2 * R1 + 1
    t = np.zeros((8, n), np.float64)
    dt = np.zeros((8, n), np.float64)
    if lib().toed_oracle_taps_frozen(float(sigma), int(R0), int(R1), _p(t), _p(dt)):
        raise ValueError("bad taps request")
    return t, dt


def _weights(w, n):
    w = np.asarray(w, np.float64)
    if w.ndim == 1:
        w = np.tile(w.reshape(1, 4), (n, 1))
    return np.ascontiguousarray(w.reshape(n, 4))


def frozen_loss(img, sigma, eidx, w, R0=None, R1=None, octants=None, outputs=False):
    """L = sum_e w_e . [x, y, theta, s] on the frozen edgel set eidx (interp
    indices) with frozen octants (decided here when None) and frozen window
    radii.  Returns (L, octants[, outputs])."""
    img = np.ascontiguousarray(img, np.float64)
    H, W = img.shape
    if R0 is None:
        R0, R1 = radius(sigma)
    eidx = np.ascontiguousarray(eidx, np.int64)
    n = len(eidx)
    oc = np.zeros(n, np.int32) if octants is None else np.ascontiguousarray(octants, np.int32).copy()
    outs = np.zeros((n, 4), np.float64)
    L = lib().toed_oracle_frozen_loss(_p(img), H, W, float(sigma), int(R0), int(R1), _p(eidx, _i64p), n,
                                      _p(_weights(w, n)), oc.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                                      _p(outs))
    return (L, oc, outs) if outputs else (L, oc)


def frozen_grad(img, sigma, eidx, w, octants, R0=None, R1=None):
    """(dL/dimg, dL/dsigma) of frozen_loss by the independent route of
    toed_grad_oracle.c (local central differences + transposed filters)."""
    img = np.ascontiguousarray(img, np.float64)
    H, W = img.shape
    if R0 is None:
        R0, R1 = radius(sigma)
    eidx = np.ascontiguousarray(eidx, np.int64)
    n = len(eidx)
    g = np.zeros((H, W), np.float64)
    gs = ctypes.c_double()
    oc = np.ascontiguousarray(octants, np.int32)
    if lib().toed_oracle_frozen_grad(_p(img), H, W, float(sigma), int(R0), int(R1), _p(eidx, _i64p), n,
                                     _p(_weights(w, n)), oc.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), _p(g),
                                     ctypes.byref(gs)):
        raise ValueError("frozen_grad failed")
    return g, gs.value


# ---- subpixel third-order correction (subpix_oracle.c) ---------------------

def subpix_taps(sigma):
    """(radius, G, G', G'') at integer offsets -r..r, r = ceil(4 sigma)."""
    G, dG, d2G = (np.zeros(129, np.float64) for _ in range(3))
    r = ctypes.c_int()
    if lib().toed_subpix_taps(float(sigma), ctypes.byref(r), _p(G), _p(dG), _p(d2G)):
        raise ValueError("bad sigma %r" % sigma)
    n = 2 * r.value + 1
    return r.value, G[:n].copy(), dG[:n].copy(), d2G[:n].copy()


def subpix_estimate_orientation(E):
    E = np.ascontiguousarray(E, np.float64)
    th = np.empty_like(E)
    lib().toed_subpix_estimate_orientation(_p(E), E.shape[0], E.shape[1], _p(th))
    return th


def subpix_conv_circ(src, kx, ky):
    src = np.ascontiguousarray(src, np.float64)
    kx = np.ascontiguousarray(kx, np.float64)
    ky = np.ascontiguousarray(ky, np.float64)
    assert len(kx) == len(ky) and len(kx) % 2 == 1
    dst = np.empty_like(src)
    lib().toed_subpix_conv_circ(_p(src), _p(dst), src.shape[0], src.shape[1], len(kx) // 2, _p(kx), _p(ky))
    return dst


def subpix_interp_cubic(src, x, y):
    src = np.ascontiguousarray(src, np.float64)
    return lib().toed_subpix_interp_cubic(_p(src), src.shape[0], src.shape[1], float(x), float(y))


def work_shape(h, w, interp_n=0):
    """Work grid of a map with interp_n > 0 (subpix_to_correction.cpp:102-104)."""
    s = 1 << int(interp_n)
    return (h - 1) * s + 1, (w - 1) * s + 1


def resize(src, dh, dw, f32=False):
    """Bilinear resize with aligned corners (subpix_to_correction.cpp:69-90), FP64."""
    src = np.ascontiguousarray(src, np.float64)
    out = np.empty((dh, dw), np.float64)
    lib().toed_subpix_resize(_p(src), src.shape[0], src.shape[1], _p(out), dh, dw, int(bool(f32)))
    return out


def subpix(E, theta=None, thresh=0.2, sigma=2.0, nbr_radius=2.0, nthreads=0, maps=False, interp_n=0,
           work_f32=False):
    """FP64 subpixel third-order correction of an edge map.  Returns dict with
    ``edges`` (N x 4: x, y, orientation, strength), ``idx`` (work-grid pixel
    i*W'+j of each edgel; the map pixel when interp_n = 0) and, with
    maps=True, ``TO_E`` / ``TO_O`` (map size).  interp_n > 0 runs the path on
    the bilinearly upsampled work grid; work_f32 rounds the resized maps to
    float, as the GPU stores them."""
    E = np.ascontiguousarray(E, np.float64)
    h, w = E.shape
    th = None if theta is None else np.ascontiguousarray(theta, np.float64)
    wh, ww = work_shape(h, w, interp_n)
    cap = wh * ww
    edges = np.empty((cap, 4), np.float64)
    idx = np.empty(cap, np.int64)
    te = np.empty((h, w), np.float64) if maps else None
    to = np.empty((h, w), np.float64) if maps else None
    n = lib().toed_subpix_oracle_n(_p(E), h, w, _p(th), float(thresh), float(sigma), float(nbr_radius),
                                   int(nthreads), int(interp_n), int(bool(work_f32)), _p(edges), cap,
                                   _p(idx, _i64p), _p(te), _p(to))
    if n < 0:
        raise ValueError("subpix oracle failed")
    res = dict(edges=edges[:n].copy(), idx=idx[:n].copy(), n=int(n))
    if maps:
        res["TO_E"], res["TO_O"] = te, to
    return res


SUBPIX_REF_BIN = os.path.join(HERE, "_ref", "subpix_ref")


def subpix_ref_available():
    return os.path.exists(SUBPIX_REF_BIN)


def run_subpix_ref(E, theta=None, thresh=0.2, sigma=2.0, nbr_radius=2.0, nthreads=1, interp_n=0):
    """Run the reference's own subpix_to_correction.cpp (compiled into
    oracle/_ref), constructed with n = interp_n.  Returns dict(edges, TO_E,
    TO_O, times=(nms, conv, correct, total) seconds)."""
    E = np.ascontiguousarray(E, np.float64)
    h, w = E.shape
    with tempfile.TemporaryDirectory() as td:
        mp = os.path.join(td, "map.f64")
        E.tofile(mp)
        op = "-"
        if theta is not None:
            op = os.path.join(td, "orient.f64")
            np.ascontiguousarray(theta, np.float64).tofile(op)
        out = os.path.join(td, "out.bin")
        subprocess.run([SUBPIX_REF_BIN, mp, op, str(h), str(w), repr(float(thresh)), repr(float(sigma)),
                        str(int(nthreads)), repr(float(nbr_radius)), out, td, str(int(interp_n))], check=True,
                       stdout=subprocess.DEVNULL)
        with open(out, "rb") as f:
            n = int(np.frombuffer(f.read(8), np.int64)[0])
            times = tuple(float(t) for t in np.frombuffer(f.read(32), np.float64))
            edges = np.frombuffer(f.read(n * 32), np.float64).reshape(n, 4).copy()
            te = np.frombuffer(f.read(h * w * 8), np.float64).reshape(h, w).copy()
            to = np.frombuffer(f.read(h * w * 8), np.float64).reshape(h, w).copy()
    return dict(edges=edges, TO_E=te, TO_O=to, times=times, n=n)


class SubpixFrozen:
    """The subpixel correction path as a function of the map E with every
    discrete decision frozen at E0 (edgel set idx, octants, the half-turn of
    each pixel's direction, sign(-Oxy), cubic stencil bases): FP64, for
    finite-difference checks of the GPU adjoint.  theta: a given orientation
    map (radians), a constant input; None estimates it from E."""

    def __init__(self, E0, idx, sigma=2.0, theta=None):
        E0 = np.ascontiguousarray(E0, np.float64)
        self.h, self.w = E0.shape
        self.sigma = float(sigma)
        self.theta = None if theta is None else np.ascontiguousarray(theta, np.float64)
        self.idx = np.ascontiguousarray(idx, np.int64)
        n, npx = len(self.idx), E0.size
        self.tneg = np.zeros(npx, np.int8)
        self.sxy = np.zeros(npx, np.int8)
        self.oct = np.zeros(n, np.int32)
        self.base = np.zeros(2 * n, np.int32)
        self.fb = np.zeros(n, np.int32)
        i8 = ctypes.POINTER(ctypes.c_byte)
        ip = ctypes.POINTER(ctypes.c_int)
        lib().toed_subpix_frozen_state(_p(E0), self.h, self.w, _p(self.theta), _p(self.idx, _i64p), n,
                                       self.tneg.ctypes.data_as(i8), self.sxy.ctypes.data_as(i8),
                                       self.oct.ctypes.data_as(ip), self.base.ctypes.data_as(ip),
                                       self.fb.ctypes.data_as(ip), self.sigma)

    def loss(self, E, w, outputs=False):
        E = np.ascontiguousarray(E, np.float64)
        w = np.ascontiguousarray(w, np.float64)
        per = 1 if w.ndim == 2 else 0
        n = len(self.idx)
        outs = np.zeros((n, 4), np.float64)
        i8 = ctypes.POINTER(ctypes.c_byte)
        ip = ctypes.POINTER(ctypes.c_int)
        L = lib().toed_subpix_frozen_loss(_p(E), self.h, self.w, _p(self.theta), self.sigma, _p(self.idx, _i64p), n,
                                          _p(w), per, self.tneg.ctypes.data_as(i8), self.sxy.ctypes.data_as(i8),
                                          self.oct.ctypes.data_as(ip), self.base.ctypes.data_as(ip),
                                          self.fb.ctypes.data_as(ip), _p(outs))
        return (L, outs) if outputs else L


CURVELET_REF_BIN = os.path.join(HERE, "_ref", "curvelet_ref_ndebug")


def curvelet_ref_available():
    return os.path.exists(CURVELET_REF_BIN)


def run_curvelet_ref(edges, H, W, types="012", rad=3.5, gap=1.5, dx=0.4, dt=15.0 / 180.0 * np.pi, token_len=1.0,
                     max_k=0.3, style=2, group=4, binary=None):
    """Run the reference's own curvelet code (curvelet/*.hpp compiled into
    oracle/_ref by oracle/Makefile) on an N x 4 list, driven as
    main_gpu_cpu.cpp:211-281 drives it.  Returns (seconds of execute(),
    {type: (chain int32 h x w, info float64 h x info_w)}); the arrays are the
    reference's column-major outputs, returned as row-major views.  The
    default binary is the -DNDEBUG build (output type 2 stops at an assert in
    the reference's own build); types 0 and 1 are the same in both."""
    e = np.ascontiguousarray(np.asarray(edges, np.float64)[:, :4]) if len(edges) else np.zeros((0, 4))
    with tempfile.TemporaryDirectory() as td:
        inp = os.path.join(td, "edges.f64")
        out = os.path.join(td, "out.bin")
        e.tofile(inp)
        subprocess.run([binary or CURVELET_REF_BIN, inp, str(len(e)), str(int(H)), str(int(W)), repr(float(rad)),
                        repr(float(gap)), repr(float(dx)), repr(float(dt)), repr(float(token_len)),
                        repr(float(max_k)), str(int(style)), str(int(group)), types, out], check=True,
                       stdout=subprocess.DEVNULL)
        b = open(out, "rb").read()
    secs = float(np.frombuffer(b, np.float64, 1, 0)[0])
    o, res = 8, {}
    while o < len(b):
        ty, h, w, iw = (int(v) for v in np.frombuffer(b, np.int32, 4, o))
        o += 16
        chain = np.frombuffer(b, np.int32, h * w, o).reshape(w, h).T.copy()
        o += 4 * h * w
        info = np.frombuffer(b, np.float64, h * iw, o).reshape(iw, h).T.copy()
        o += 8 * h * iw
        res[ty] = (chain, info)
    return secs, res
