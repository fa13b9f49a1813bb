"""TOED detection throughput on one MI355X.

One JSON line on stdout.  Headline: plain single-image TOED on 2160x3840 u8
frames (sigma = 2, threshold 2), Mpixel/s with the frame resident in HBM.
Also reported ("forms"): the synchronous per-call form, the batched form (200
seeded scenes of mixed size and sigma in one plan, replayed as a hipGraph)
against a loop of plain launches over the same images, the differentiable
forward+backward form, the subpixel third-order correction of 32 Bezier edge
maps of 2048x2048, kernel durations from HIP events on the library's stream,
the VALU/HBM roofline of the dominant kernel, the curvelet formation stage on
the frame's edgel list, and the reference CPU paths timed on this host.

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--workload all|plain|batch|fwdbwd|subpix]
"""
import argparse
import gc
import importlib
import json
import os
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
PKG_NAME = "third-order-edge-detector_amd"

METRIC = "TOED Mpixel/s at 3840x2160: batched & fwd+bwd vs plain, same GPU; % HBM peak"
PEAK_FP32_TFLOPS = 157.3
PEAK_HBM_GBS = 8000.0
# HBM bytes per launch of the plain-frame tile kernel (2160x3840 u8, sigma 2), from
# rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE runs of this workload, FETCH_SIZE scaled by
# 2.0 (what the counter reads for a 256 MiB copy); the summary is kept under profiles/.
TRAFFIC_TILES_PLAIN_BYTES = 2.0 * 24211.2 * 1024 + 19422.8 * 1024
BATCH_SIZES = [(480, 640), (480, 752), (720, 1280), (1080, 1920), (2160, 3840)]
BATCH_SIGMAS = [1.0, 1.5, 2.0]


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def dense_flops_per_px(sigma, pkg):
    """Minimal useful dense work of the tile kernel per input pixel: Ix and Iy
    at the four interp phases, separable: 10*W1 + 2*W0 FMAs (2 flops each)."""
    r0, r1 = pkg.sigma_radius(sigma)
    w0, w1 = 2 * r0 + 1, 2 * r1 + 1
    return 2.0 * (10 * w1 + 2 * w0)


def time_loop(torch, fn, steps, warmup):
    # Python's cyclic garbage collector is paused while a loop is timed (as
    # timeit does): a collection pass over the objects this process holds is
    # host noise unrelated to the work timed
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    was = gc.isenabled()
    gc.disable()
    try:
        t0 = time.perf_counter()
        for _ in range(steps):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / steps
    finally:
        if was:
            gc.enable()


def bench_plain(torch, pkg, scene, steps, warmup):
    H, W = 2160, 3840
    img_np = scene.shapes_scene(2160 * 3840, H, W)
    img = torch.from_numpy(img_np).cuda()
    det = pkg.Detector(0)
    cap = 4 * H * W
    edges = torch.empty((cap, 4), dtype=torch.float32, device="cuda")
    counts = torch.zeros(2, dtype=torch.int32, device="cuda")
    sec = time_loop(torch, lambda: det.detect_into(img, edges, counts), steps, warmup)
    n, flags = (int(v) for v in counts.cpu())
    assert flags == 0, flags
    edges_np = edges[:n].cpu().numpy()
    # per-kernel durations (HIP events on the library stream), synchronous form
    tiles, fin = [], []
    t0 = time.perf_counter()
    for _ in range(steps):
        det.detect_sync_into(img, edges)
        a, b = det.last_times()
        tiles.append(a)
        fin.append(b)
    sync_sec = (time.perf_counter() - t0) / steps
    # the optional orientation at the subpixel location (MATLAB QaD mode)
    det.set_orientation_mode(1)
    qad_sec = time_loop(torch, lambda: det.detect_into(img, edges, counts), max(3, steps // 4), 1)
    det.set_orientation_mode(0)
    det.close()
    return dict(H=H, W=W, sec=sec, n=n, ms_tiles=float(np.mean(tiles)), ms_final=float(np.mean(fin)),
                sync_sec=sync_sec, img_np=img_np, edges_np=edges_np, qad_sec=qad_sec)


def batch_images(scene, n=200, seed=2024):
    rng = np.random.default_rng(seed)
    shapes = [BATCH_SIZES[k] for k in rng.integers(0, len(BATCH_SIZES), n)]
    sigmas = [BATCH_SIGMAS[k] for k in rng.integers(0, len(BATCH_SIGMAS), n)]
    imgs = [scene.shapes_scene(seed * 1000 + m, H, W) for m, (H, W) in enumerate(shapes)]
    return imgs, sigmas


def bench_batch(torch, pkg, scene, steps, warmup, n_images):
    imgs, sigmas = batch_images(scene, n_images)
    dev = [torch.from_numpy(im).cuda() for im in imgs]
    px = float(sum(im.size for im in imgs))
    det = pkg.Detector(0, staging_fraction=0.25)
    caps = [max(1024, (4 * im.size) // 8) for im in imgs]
    # the project's autotuner picks tile rows and the launch grid on these images
    tune = det.autotune_batch(dev, sigmas, capacities=caps, reps=3)
    best = tune["best"]
    b = pkg.BatchDetector(det, dev, sigmas, capacities=caps, persistent=best["persistent"],
                          wg_per_cu=best["wg_per_cu"])
    sec_graph = time_loop(torch, lambda: b.run(graph=True), steps, warmup)
    counts = b.counts.cpu().numpy()
    assert not b.flags.cpu().numpy().any()
    b_plain = pkg.BatchDetector(det, dev, sigmas, capacities=caps, persistent=False)
    sec_batch_plain = time_loop(torch, lambda: b_plain.run(graph=False), steps, warmup)
    # loop of plain single-image launches over the same images and outputs
    cnt = torch.zeros((len(dev), 2), dtype=torch.int32, device="cuda")
    views = [b.edges[int(b.offsets[m]):int(b.offsets[m]) + caps[m]] for m in range(len(dev))]

    def loop():
        for m in range(len(dev)):
            det.detect_into(dev[m], views[m], cnt[m], sigmas[m])

    sec_loop = time_loop(torch, loop, max(2, steps // 4), 1)
    loop_counts = cnt[:, 0].cpu().numpy()
    assert np.array_equal(loop_counts, counts), "batched and per-image counts differ"
    b.close()
    b_plain.close()
    det.close()
    return dict(px=px, n_images=len(imgs), sec_graph=sec_graph, sec_batch_plain=sec_batch_plain, sec_loop=sec_loop,
                edgels=int(counts.sum()), sigmas=sigmas, shapes=[im.shape for im in imgs], tune=tune)


def bench_fwd_bwd(torch, pkg, scene, steps, warmup, B=64, H=512, W=512, sigma=1.5):
    """Differentiable form: B float32 scenes, loss sum(a x + b y + c theta + d s)
    with (a, b, c, d) ~ N(0, 1) fixed from the seed; gradients w.r.t. every pixel
    and sigma.  Reports forward-only and forward+backward throughput."""
    imgs = np.stack([scene.shapes_scene(9000 + b, H, W) for b in range(B)]).astype(np.float32)
    x = torch.from_numpy(imgs).cuda().requires_grad_(True)
    # sigma is a parameter in device memory, as a training loop holds it
    sig = torch.tensor(sigma, dtype=torch.float64, device="cuda", requires_grad=True)
    w = torch.tensor(scene.loss_weights(12345), dtype=torch.float32, device="cuda")
    det = pkg.Detector(0, staging_fraction=0.25)
    state = {}

    def fwd():
        with torch.no_grad():
            e, off = pkg.detect_differentiable(x, sig, detector=det)
        state["n"] = e.shape[0]

    def fwd_bwd():
        x.grad = None
        sig.grad = None
        e, off = pkg.detect_differentiable(x, sig, detector=det)
        (e * w).sum().backward()

    t_f = time_loop(torch, fwd, steps, warmup)
    t_fb = time_loop(torch, fwd_bwd, steps, warmup)
    # kernel time of the same plan, HIP events on the library's stream (torch's
    # current stream): its forward launch sequence, and its backward
    autograd = importlib.import_module(PKG_NAME + ".autograd")
    plan = next(reversed(autograd._PLANS.values()))
    plan.point_at(x.detach())
    bd = plan.bd
    bd.run()
    counts, _ = bd.counts_flags()
    offs = np.concatenate([[0], np.cumsum(counts.astype(np.int64))])
    g = (torch.ones((max(int(offs[-1]), 1), 4), device="cuda") * w).contiguous()
    gimg = torch.zeros((B, H, W), dtype=torch.float32, device="cuda")
    gsig = torch.zeros(1, dtype=torch.float64, device="cuda")
    gp = [g.data_ptr() + 16 * int(offs[b]) for b in range(B)]
    ip = [gimg.data_ptr() + 4 * H * W * b for b in range(B)]
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
    kf, kb = [], []
    for _ in range(max(3, steps)):
        ev[0].record()
        bd.run()
        ev[1].record()
        bd.backward_ptrs(gp, ip, [W] * B, gsig.data_ptr())
        ev[2].record()
        torch.cuda.synchronize()
        kf.append(ev[0].elapsed_time(ev[1]))
        kb.append(ev[1].elapsed_time(ev[2]))
    det.close()
    return dict(px=float(B * H * W), t_f=t_f, t_fb=t_fb, n=state["n"], k_f=float(np.median(kf)) * 1e-3,
                k_b=float(np.median(kb)) * 1e-3)


def bench_subpix(torch, pkg, scene, steps, warmup, B=32, S=2048):
    """Subpixel third-order correction on B float32 Bezier edge maps of S x S
    (40 curves, thresh 0.1, sigma 2, orientation estimated, isolation radius
    2) in one plan; then the differentiable form (autograd wrapper, loss
    sum(a x + b y + c theta + d s) with (a, b, c, d) ~ N(0, 1) fixed from the
    seed, gradient w.r.t. every map pixel): forward alone and forward+backward."""
    x = torch.from_numpy(np.stack([scene.bezier_map(5000 + b, S, S, 40) for b in range(B)])).cuda()
    det = pkg.Detector(0)
    cap = S * S // 8
    sp = pkg.SubpixBatch(det, [x[b] for b in range(B)], None, 0.1, 2.0, 2.0, capacities=[cap] * B)
    sec = time_loop(torch, sp.run, steps, warmup)
    counts = sp.counts.cpu().numpy()
    assert not sp.flags.cpu().numpy().any()
    edges0 = sp.results()[0].cpu().numpy()
    sp.close()
    xg = x.clone().requires_grad_(True)
    w = torch.tensor(scene.loss_weights(12345), dtype=torch.float32, device="cuda")

    def fwd():
        with torch.no_grad():
            pkg.subpix_differentiable(xg, 0.1, 2.0, 2.0, detector=det, capacity=cap)

    def fwd_bwd():
        xg.grad = None
        e, off = pkg.subpix_differentiable(xg, 0.1, 2.0, 2.0, detector=det, capacity=cap)
        (e * w).sum().backward()

    t_f = time_loop(torch, fwd, steps, warmup)
    t_fb = time_loop(torch, fwd_bwd, steps, warmup)
    map0 = x[0].cpu().numpy()
    det.close()
    return dict(px=float(B * S * S), sec=sec, edgels=int(counts.sum()), map0=map0, edges0=edges0, B=B, S=S,
                t_f=t_f, t_fb=t_fb)


CPU_SAMPLE_SEC = 10.0   # CPU baseline: run the sample until this much measured time
CPU_MAX_RUNS = 25


def cpu_baseline_subpix(map_np, gpu_edges=None):
    """The reference's subpix_to_correction.cpp (FP64, OpenMP) on one map of the
    subpix workload; the oracle port when the reference harness is not built.
    Its first run's list is also the checker of the GPU list of that map."""
    import oracle
    cores = host_cpu()["threads_used"]
    H, W = map_np.shape
    sec, runs = 0.0, 0
    first = None
    while runs < CPU_MAX_RUNS and sec < CPU_SAMPLE_SEC:   # a bounded sample: about 10 s of CPU work
        if oracle.subpix_ref_available():
            r = oracle.run_subpix_ref(map_np.astype(np.float64), None, 0.1, 2.0, 2.0, nthreads=cores)
            sec, kind, edges = sec + r["times"][3], "reference", r["edges"]
        else:
            t0 = time.perf_counter()
            edges = oracle.subpix(map_np.astype(np.float64), None, 0.1, 2.0, 2.0, nthreads=cores)["edges"]
            sec, kind = sec + time.perf_counter() - t0, "port"
        if first is None:
            first = edges
        runs += 1
    res = {"value": round(runs * H * W / sec / 1e6, 4), "unit": "Mpixel/s", "cores": cores, "kind": kind,
           "sample": "%d runs on one %dx%d Bezier map, FP64, orientation estimate through the edgel list "
                     "(%.2f s in total)" % (runs, H, W, sec)}
    return res, (accuracy(gpu_edges, first) if gpu_edges is not None else None)


def bench_curvelets(pkg, rows, H, W, threads, reps=5, ref_band=540):
    """Curvelet formation on the headline frame's GPU list (widened to double,
    as the reference's driver hands its FP64 list over, main_gpu_cpu.cpp:
    211-281) with the drivers' settings: the library on `threads` host
    threads (best of `reps` runs of execute()), and -- when the reference's
    own curvelet code is built under oracle/_ref -- that code on a bounded
    sample: the edgels of the frame's top `ref_band` rows, timed around its
    execute(), with the two outputs compared bit for bit on the sample."""
    e = np.ascontiguousarray(rows, np.float64)
    best = None
    for _ in range(reps):
        cf = pkg.CurveletFormation(e, H, W)
        t0 = time.perf_counter()
        cf.execute(threads)
        dt = time.perf_counter() - t0
        best = dt if best is None else min(best, dt)
    h, _ = cf.get_output_size(0)
    res = {"edgels": int(len(e)), "curvelets": int(h), "ms": round(best * 1e3, 2),
           "edgels_s": round(len(e) / best, 1), "threads": threads,
           "settings": "nrad 3.5, gap 1.5, dx 0.4, dt 15 deg, max_k 0.3, style 2, groups of 4; all four stages"}
    import oracle
    if oracle.curvelet_ref_available():
        band = e[e[:, 1] < ref_band]
        secs, ref = oracle.run_curvelet_ref(band, ref_band, W, types="01",
                                            binary=os.path.join(ROOT, "oracle", "_ref", "curvelet_ref"))
        cb = pkg.CurveletFormation(band, ref_band, W)
        t0 = time.perf_counter()
        cb.execute(threads)
        ours_band = time.perf_counter() - t0
        same = True
        for t in (0, 1):
            c, i = cb.get_output_arrary(t)
            same = same and np.array_equal(c, ref[t][0]) and np.array_equal(i.view(np.int64), ref[t][1].view(np.int64))
        res["cpu_reference"] = {"edgels_s": round(len(band) / secs, 1), "threads": 1, "kind": "reference",
                                "sample": "%d edgels of the frame's top %d rows, one run of execute() (%.2f s)"
                                          % (len(band), ref_band, secs),
                                "library_on_sample_ms": round(ours_band * 1e3, 2),
                                "bit_exact_on_sample": bool(same)}
        res["vs_reference"] = round(res["edgels_s"] / res["cpu_reference"]["edgels_s"], 1)
    return res


def accuracy(gpu, ref):
    """Errors of the GPU edgel list against the FP64 CPU list of the same frame:
    edgels matched by position (nearest within 1e-2 px), then the largest
    position, orientation and relative strength differences over the matches."""
    from scipy.spatial import cKDTree
    ref = np.asarray(ref, np.float64)
    gpu = np.asarray(gpu, np.float64)
    out = {"edgels_gpu": int(len(gpu)), "edgels_cpu": int(len(ref))}
    if len(gpu) == 0 or len(ref) == 0:
        return out
    d, j = cKDTree(ref[:, :2]).query(gpu[:, :2], distance_upper_bound=1e-2)
    ok = np.isfinite(d)
    g, r = gpu[ok], ref[j[ok]]
    dth = np.abs((g[:, 2] - r[:, 2] + np.pi) % (2 * np.pi) - np.pi)
    out.update({"matched": int(ok.sum()),
                "max_pos_err_px": float(np.abs(g[:, :2] - r[:, :2]).max()),
                "max_orient_err_rad": float(dth.max()),
                "max_strength_rel_err": float((np.abs(g[:, 3] - r[:, 3]) / np.maximum(np.abs(r[:, 3]), 1e-12)).max()),
                "tolerances": "1e-3 px, 1e-3 rad, 1e-4 relative"})
    return out


def host_cpu():
    """The host's CPU as lscpu reports it, and the threads this process may use:
    the CPUs in its affinity mask, limited by OMP_NUM_THREADS where the job
    sets one (a shared host may allot each GPU a share of its cores)."""
    info = {}
    try:
        out = subprocess.run(["lscpu"], stdout=subprocess.PIPE, stderr=subprocess.DEVNULL, text=True).stdout
        kv = dict(l.split(":", 1) for l in out.splitlines() if ":" in l)
        kv = {k.strip(): v.strip() for k, v in kv.items()}
        info["model"] = kv.get("Model name", "?")
        sockets = int(kv.get("Socket(s)", "1") or 1)
        cps = int(kv.get("Core(s) per socket", "0") or 0)
        info["physical_cores"] = sockets * cps if cps else None
        info["logical_cpus"] = int(kv.get("CPU(s)", "0") or 0)
    except (OSError, ValueError):
        info["model"] = "?"
    avail = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    omp = os.environ.get("OMP_NUM_THREADS")
    n = avail
    if omp and omp.isdigit() and int(omp) > 0:
        n = min(n, int(omp))
    info["cpus_available"] = avail
    info["threads_used"] = max(1, n)
    return info


def _time_ref(img_np, threads, budget, max_runs, env):
    """Repeated runs of the reference CPU path (or the oracle port) on one
    image until about `budget` seconds of measured time: (Mpixel/s, kind, runs,
    seconds, first list).  The reference's runs are timed with its own
    omp_get_wtime intervals around convolve_img and non_maximum_suppresion
    (cpu_toed.cpp:171/359, 390/499)."""
    import oracle
    H, W = img_np.shape
    sec, runs, first, kind = 0.0, 0, None, "port"
    while runs < max_runs and (runs == 0 or sec < budget):
        if oracle.ref_available():
            edges, tc, tn = oracle.run_ref(img_np, nthreads=threads, env=env)
            kind = "reference"
            sec += tc + tn
        else:
            t0 = time.perf_counter()
            edges = oracle.detect(img_np.astype(np.float64), 2.0, nthreads=threads)["edges"]
            sec += time.perf_counter() - t0
        if first is None:
            first = edges
        runs += 1
    return runs * H * W / sec / 1e6, kind, runs, sec, first


def cpu_baseline(img_np, gpu_edges=None, scene=None):
    """The reference CPU path (cpu_toed.cpp, FP64, OpenMP, schedule(dynamic))
    on this host, as the baseline protocol states it: the 480x752 scene at one
    thread and at N threads, and one frame of the 2160x3840 workload at N
    threads (the reported value), N = the threads available to this process,
    OMP_PROC_BIND=close; the oracle port when the reference harness is not
    built.  Its first 2160x3840 list is also the checker of the GPU list of the
    same frame (``accuracy``)."""
    host = host_cpu()
    n = host["threads_used"]
    env = {"OMP_PROC_BIND": "close", "OMP_PLACES": "cores"}
    H, W = img_np.shape
    v, kind, runs, sec, first = _time_ref(img_np, n, CPU_SAMPLE_SEC, CPU_MAX_RUNS, env)
    res = {"value": round(v, 4), "unit": "Mpixel/s", "cores": n, "kind": kind,
           "sample": "%d runs on one %dx%d frame of the same workload at %d threads (OMP_PROC_BIND=close), FP64, "
                     "timed at the reference's omp_get_wtime boundaries of convolve_img + non_maximum_suppresion "
                     "(%.2f s in total)" % (runs, H, W, n, sec),
           "host": host}
    if scene is not None:
        c1 = scene.shapes_scene(480 * 752, 480, 752)
        v1, _, r1, s1, _ = _time_ref(c1, 1, 4.0, 3, env)
        vn, _, rn, sn, _ = _time_ref(c1, n, 2.0, 20, env)
        res["c1_480x752"] = {"threads_1": {"mpx_s": round(v1, 4), "runs": r1, "sec": round(s1, 3)},
                             "threads_n": {"threads": n, "mpx_s": round(vn, 4), "runs": rn, "sec": round(sn, 3)},
                             "speedup": round(vn / v1, 2)}
    acc = accuracy(gpu_edges, first) if gpu_edges is not None else None
    return res, acc


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="all", choices=["all", "plain", "batch", "fwdbwd", "subpix"])
    ap.add_argument("--batch-images", type=int, default=200)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (each image is independent; "
              "it would shard by image, one process per GPU). Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)
    import torch
    pkg = importlib.import_module(PKG_NAME)
    scene = importlib.import_module(PKG_NAME + ".scene")
    torch.cuda.init()
    forms = {}
    out = {"metric": METRIC, "unit": "Mpixel/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
           "data": "synthetic (seeded shape scenes)"}
    plain = None
    if args.workload in ("all", "plain"):
        log("plain 2160x3840 ...")
        plain = bench_plain(torch, pkg, scene, args.steps, args.warmup)
        px = plain["H"] * plain["W"]
        mpx = px / plain["sec"] / 1e6
        forms["plain"] = {"mpx_s": round(mpx, 1), "ms": round(plain["sec"] * 1e3, 4), "edgels": plain["n"],
                          "edgels_s": round(plain["n"] / plain["sec"], 1)}
        forms["plain_sync"] = {"mpx_s": round(px / plain["sync_sec"] / 1e6, 1),
                               "ms": round(plain["sync_sec"] * 1e3, 4)}
        forms["kernels_ms"] = {"tiles": round(plain["ms_tiles"], 4), "scan_finalize": round(plain["ms_final"], 4)}
        forms["plain_orientation_at_subpixel"] = {"mpx_s": round(px / plain["qad_sec"] / 1e6, 1),
                                                  "ms": round(plain["qad_sec"] * 1e3, 4)}
        flops = dense_flops_per_px(2.0, pkg) * px
        ach = flops / (plain["ms_tiles"] * 1e-3) / 1e12
        out["roofline"] = {"bound": "valu", "achieved": round(ach, 2), "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                           "frac": round(ach / PEAK_FP32_TFLOPS, 4), "traffic": round(TRAFFIC_TILES_PLAIN_BYTES),
                           "traffic_unit": "bytes per launch (rocprofv3 FETCH_SIZE x 2.0 + WRITE_SIZE)",
                           "kernel": "toed_tiles_kernel<uint8_t,16>",
                           "algorithmic": "2*(10*W1+2*W0) flop/px = %d flop/px at sigma 2, %d px/launch"
                                          % (dense_flops_per_px(2.0, pkg), px)}
        # the whole frame against the same roof, with the per-edgel term of the
        # minimal useful work (the seven higher derivatives at each edgel,
        # 4*W1^2 + 7*W1 FMAs, which the finalize kernel evaluates)
        r0, r1 = pkg.sigma_radius(2.0)
        w1 = 2 * r1 + 1
        fflops = flops + 2.0 * (4 * w1 * w1 + 7 * w1) * plain["n"]
        fach = fflops / plain["sec"] / 1e12
        out["roofline_frame"] = {"achieved": round(fach, 2), "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                                 "frac": round(fach / PEAK_FP32_TFLOPS, 4),
                                 "algorithmic": "2*(10*W1+2*W0) flop/px + 2*(4*W1^2+7*W1) flop/edgel = %.3f GFLOP "
                                                "per frame (%d edgels), whole launch sequence" % (fflops / 1e9,
                                                                                                 plain["n"])}
        hb = (px * 1 + 16 * plain["n"] + 4) / plain["sec"] / 1e9
        out["hbm"] = {"achieved": round(hb, 2), "peak": PEAK_HBM_GBS, "unit": "GB/s", "frac": round(hb / PEAK_HBM_GBS, 5),
                      "algorithmic_bytes": "1 B/px in + 16 B/edgel + 4 B count"}
        out["value"] = round(mpx, 1)
        out["ms_per_step"] = round(plain["sec"] * 1e3, 4)
        out["config"] = {"workload": "plain TOED: one 2160x3840 u8 synthetic frame per step (96 shapes), "
                                     "sigma 2 (R0 8 / R1 9), threshold 2, border 10, FP32 + FP64 per edgel",
                         "global_batch": 1, "H": 2160, "W": 3840}
        log("plain: %.1f Mpx/s, %.3f ms/frame, %d edgels, tiles %.3f ms" % (mpx, plain["sec"] * 1e3, plain["n"],
                                                                           plain["ms_tiles"]))
    if args.workload in ("all", "batch"):
        log("batch of %d ..." % args.batch_images)
        bt = bench_batch(torch, pkg, scene, max(3, args.steps // 4), 1, args.batch_images)
        forms["batched_graph"] = {"mpx_s": round(bt["px"] / bt["sec_graph"] / 1e6, 1),
                                  "ms": round(bt["sec_graph"] * 1e3, 3),
                                  "edgels_s": round(bt["edgels"] / bt["sec_graph"], 1)}
        forms["batched_plain_grid"] = {"mpx_s": round(bt["px"] / bt["sec_batch_plain"] / 1e6, 1)}
        forms["batch_loop_of_plain"] = {"mpx_s": round(bt["px"] / bt["sec_loop"] / 1e6, 1),
                                        "ms": round(bt["sec_loop"] * 1e3, 3)}
        forms["batch_config"] = {"images": bt["n_images"], "mpx": round(bt["px"] / 1e6, 2), "edgels": bt["edgels"]}
        # the images' sigma differ (windows of 9/11, 13/15, 17/19 taps): algorithmic work weighted per image
        bflops = sum(float(h * w) * dense_flops_per_px(sg, pkg) for (h, w), sg in zip(bt["shapes"], bt["sigmas"]))
        ach_b = bflops / bt["sec_graph"] / 1e12
        forms["batched_graph"]["roofline"] = {
            "bound": "valu", "achieved": round(ach_b, 2), "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
            "frac": round(ach_b / PEAK_FP32_TFLOPS, 4),
            "algorithmic": "sum over images of H*W*2*(10*W1+2*W0) flop (sigma-weighted: 256/352/448 flop/px "
                           "at sigma 1/1.5/2), %.3g GFLOP per batch, whole launch sequence timed" % (bflops / 1e9)}
        forms["batched_graph"]["vs_loop_of_plain"] = round(bt["sec_loop"] / bt["sec_graph"], 3)
        forms["batched_graph"]["hbm_note"] = (
            "the fused path reads each pixel once and is bound by FP32 VALU work (about 300-450 flop per byte "
            "against a ridge of 20): it cannot reach 50% of HBM peak without adding traffic")
        tb = bt["tune"]["best"]
        forms["batch_autotune"] = {"tile_rows": tb["tile_rows"], "persistent": tb["persistent"],
                                   "wg_per_cu": tb["wg_per_cu"], "waves_per_cu": 4 * tb["wg_per_cu"],
                                   "ms": round(tb["ms"], 3), "configs_timed": len(bt["tune"]["results"])}
        if "value" not in out:
            out["value"] = forms["batched_graph"]["mpx_s"]
            out["ms_per_step"] = round(bt["sec_graph"] * 1e3, 3)
            out["config"] = {"workload": "batched TOED: 200 mixed-size scenes, mixed sigma", "global_batch": 200}
        log("batch: graph %.1f Mpx/s, loop-of-plain %.1f Mpx/s; autotuned: %d tile rows, %s (%d configurations)"
            % (forms["batched_graph"]["mpx_s"], forms["batch_loop_of_plain"]["mpx_s"], tb["tile_rows"],
               "persistent, %d workgroups per CU" % tb["wg_per_cu"] if tb["persistent"] else "one workgroup per tile",
               len(bt["tune"]["results"])))
    if args.workload in ("all", "fwdbwd"):
        log("forward+backward 64 x 512x512 ...")
        fb = bench_fwd_bwd(torch, pkg, scene, max(3, args.steps // 4), 1)
        forms["fwd_only_64x512x512"] = {"mpx_s": round(fb["px"] / fb["t_f"] / 1e6, 1), "ms": round(fb["t_f"] * 1e3, 3)}
        forms["fwd_bwd_64x512x512"] = {"mpx_s": round(fb["px"] / fb["t_fb"] / 1e6, 1), "ms": round(fb["t_fb"] * 1e3, 3),
                                       "edgels": fb["n"], "ratio_to_fwd": round(fb["t_fb"] / fb["t_f"], 3),
                                       "kernels_ms": {"forward": round(fb["k_f"] * 1e3, 3),
                                                      "backward": round(fb["k_b"] * 1e3, 3)},
                                       "ratio_kernels": round((fb["k_f"] + fb["k_b"]) / fb["k_f"], 3)}
        if "value" not in out:
            out["value"] = forms["fwd_bwd_64x512x512"]["mpx_s"]
            out["ms_per_step"] = round(fb["t_fb"] * 1e3, 3)
            out["config"] = {"workload": "differentiable TOED fwd+bwd, 64 x 512x512 f32, sigma 1.5", "global_batch": 64}
        log("fwd+bwd: %.1f Mpx/s (fwd %.1f), ratio %.2f wall clock, %.2f kernels (fwd %.3f ms, bwd %.3f ms)"
            % (forms["fwd_bwd_64x512x512"]["mpx_s"], forms["fwd_only_64x512x512"]["mpx_s"], fb["t_fb"] / fb["t_f"],
               (fb["k_f"] + fb["k_b"]) / fb["k_f"], fb["k_f"] * 1e3, fb["k_b"] * 1e3))
    if args.workload in ("all", "subpix"):
        log("subpix 32 x 2048x2048 ...")
        sp = bench_subpix(torch, pkg, scene, max(3, args.steps // 4), 1)
        forms["subpix_32x2048x2048"] = {"mpx_s": round(sp["px"] / sp["sec"] / 1e6, 1), "ms": round(sp["sec"] * 1e3, 3),
                                        "edgels": sp["edgels"], "edgels_s": round(sp["edgels"] / sp["sec"], 1)}
        forms["subpix_fwd_only_differentiable"] = {"mpx_s": round(sp["px"] / sp["t_f"] / 1e6, 1),
                                                   "ms": round(sp["t_f"] * 1e3, 3)}
        forms["subpix_fwd_bwd_32x2048x2048"] = {"mpx_s": round(sp["px"] / sp["t_fb"] / 1e6, 1),
                                                "ms": round(sp["t_fb"] * 1e3, 3),
                                                "ratio_to_fwd": round(sp["t_fb"] / sp["t_f"], 3),
                                                "ratio_to_plain_fwd": round(sp["t_fb"] / sp["sec"], 3)}
        if not args.no_cpu:
            log("subpix cpu baseline ...")
            forms["subpix_32x2048x2048"]["cpu_baseline"], acc = cpu_baseline_subpix(sp["map0"], sp["edges0"])
            forms["subpix_32x2048x2048"]["accuracy_vs_cpu_fp64"] = acc
            log("subpix accuracy vs the FP64 CPU path (map 0): %s" % json.dumps(acc))
        if "value" not in out:
            out["value"] = forms["subpix_32x2048x2048"]["mpx_s"]
            out["ms_per_step"] = round(sp["sec"] * 1e3, 3)
            out["config"] = {"workload": "subpixel TO correction, 32 x 2048x2048 f32 Bezier maps, thresh 0.1, "
                                         "sigma 2, orientation estimated", "global_batch": 32}
        log("subpix: %.1f Mpx/s, %.3f ms per batch" % (forms["subpix_32x2048x2048"]["mpx_s"], sp["sec"] * 1e3))
    if plain is not None and args.workload == "all":
        log("curvelets on the 2160x3840 list ...")
        threads = host_cpu()["threads_used"]
        cv = bench_curvelets(pkg, plain["edges_np"], plain["H"], plain["W"], threads)
        forms["curvelets_2160x3840"] = cv
        log("curvelets: %d edgels -> %d curvelets in %.1f ms on %d threads%s"
            % (cv["edgels"], cv["curvelets"], cv["ms"], threads,
               "" if "cpu_reference" not in cv else ", reference %.0f edgels/s (x%.1f), bit-exact %s"
               % (cv["cpu_reference"]["edgels_s"], cv["vs_reference"], cv["cpu_reference"]["bit_exact_on_sample"])))
    out["forms"] = forms
    if plain is not None and not args.no_cpu:
        log("cpu baseline ...")
        out["cpu_baseline"], acc = cpu_baseline(plain["img_np"], plain["edges_np"], scene)
        out["accuracy_vs_cpu_fp64"] = acc
        log("accuracy vs the FP64 CPU path: %s" % json.dumps(acc))
    print(json.dumps(out))


if __name__ == "__main__":
    main()
