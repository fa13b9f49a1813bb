"""Write the golden vectors of the curvelet formation stage.

Runs the reference's own curvelet code -- curvelet/*.hpp (header-only)
compiled from /root/reference by oracle/Makefile into
oracle/_ref/curvelet_ref_ndebug, driven as main_gpu_cpu.cpp:211-281 drives
it -- on the detection fixtures' edgel lists (tests/golden/toed_ref_*.npz, the
reference's FP64 CPU lists) and on a seeded list of random edgels with
positions on and off the image, coinciding edgels and a cluster closer than
2 dx.  Cases: the drivers' settings (style 2, groups of 4) on every list,
the other three grouping styles and other group sizes on two of them, and
one case with every other parameter changed.

Fixtures:
  curvelet_ref.json                  per case and output type 0 / 1 / 2: the
                                     output shape and the SHA-256 of the
                                     column-major chain (int32) and info
                                     (float64) bytes, little endian
  curvelet_ref_scene_64x96_s1.npz    the full outputs of the smallest case

Output type 2 is taken from the -DNDEBUG build: the reference's own build
(no NDEBUG) stops at the assert in curvelet::set_output for that type, which
the same script checks with oracle/_ref/curvelet_ref.

Run from the repo root:  python tests/golden/make_curvelet_golden.py
"""
import hashlib
import json
import os
import subprocess
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
import oracle  # noqa: E402

OUT = os.path.dirname(os.path.abspath(__file__))
LISTS = ["scene_64x96_s1", "scene_128x160_s2", "img_t_less_pgm", "scene_480x752_s3", "img_2018_pgm",
         "img_euroc_png"]
DEFAULT = dict(rad=3.5, gap=1.5, dx=0.4, dt=15.0 / 180.0 * np.pi, token_len=1.0, max_k=0.3, style=2, group=4)
VARIANTS = {
    "style0": dict(style=0), "style1": dict(style=1), "style3": dict(style=3),
    "group3": dict(group=3), "group7": dict(group=7),
    "params": dict(rad=5.0, gap=2.0, dx=0.3, dt=10.0 / 180.0 * np.pi, token_len=1.5, max_k=0.5, group=5),
}


def random_list(seed=7, H=96, W=128, n=1500):
    """Random edgels: positions on the image and up to 1.5 px outside it,
    orientations in (-pi, 2 pi), 30 exact duplicates of other edgels and a
    tight cluster of 12 (pairs closer than 2 dx)."""
    rng = np.random.default_rng(seed)
    e = np.empty((n, 4))
    e[:, 0] = rng.uniform(-1.5, W + 1.5, n)
    e[:, 1] = rng.uniform(-1.5, H + 1.5, n)
    e[:, 2] = rng.uniform(-np.pi, 2 * np.pi, n)
    e[:, 3] = rng.uniform(2.0, 40.0, n)
    dup = rng.choice(n - 30, 30, replace=False)
    e[n - 30:] = e[dup]
    e[:12, 0] = 50.0 + rng.uniform(-0.3, 0.3, 12)
    e[:12, 1] = 40.0 + rng.uniform(-0.3, 0.3, 12)
    return e, H, W


def cases():
    for name in LISTS:
        d = np.load(os.path.join(OUT, "toed_ref_%s.npz" % name))
        H, W = d["img"].shape
        yield name, d["edges"], H, W, {}
        if name in ("scene_128x160_s2", "img_t_less_pgm"):
            for vname, v in VARIANTS.items():
                yield "%s/%s" % (name, vname), d["edges"], H, W, v
    e, H, W = random_list()
    yield "random_96x128_s7", e, H, W, {}
    yield "random_96x128_s7/style0", e, H, W, VARIANTS["style0"]
    yield "empty_32x32", np.zeros((0, 4)), 32, 32, {}


def digest(a):
    # column-major bytes, as the reference's arrays hold them
    return hashlib.sha256(np.asfortranarray(a).tobytes(order="F")).hexdigest()


def main():
    oracle.build(ref=True)
    table = {}
    for name, e, H, W, v in cases():
        p = dict(DEFAULT, **v)
        _, res = oracle.run_curvelet_ref(e, H, W, types="012", **p)
        table[name] = {"H": int(H), "W": int(W), "n_edgels": int(len(e)), "params": p}
        for ty, (chain, info) in res.items():
            table[name][str(ty)] = {"shape": [int(chain.shape[0]), int(chain.shape[1]), int(info.shape[1])],
                                    "chain_sha256": digest(chain), "info_sha256": digest(info)}
        if name == "scene_64x96_s1":
            np.savez_compressed(os.path.join(OUT, "curvelet_ref_scene_64x96_s1.npz"),
                                **{"chain%d" % t: res[t][0] for t in res}, **{"info%d" % t: res[t][1] for t in res})
        print(name, {t: table[name][str(t)]["shape"] for t in range(3)})
    # the reference's own build (asserts on) stops on output type 2
    d = np.load(os.path.join(OUT, "toed_ref_scene_64x96_s1.npz"))
    try:
        oracle.run_curvelet_ref(d["edges"], *d["img"].shape, types="2",
                                binary=os.path.join(ROOT, "oracle", "_ref", "curvelet_ref"))
        aborts = False
    except subprocess.CalledProcessError:
        aborts = True
    table["_reference_build_type2_aborts"] = aborts
    with open(os.path.join(OUT, "curvelet_ref.json"), "w") as f:
        json.dump(table, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
