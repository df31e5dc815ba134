#!/usr/bin/env python3
"""Benchmark of line2Dup's Detector::match on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload copies1024|bank36k|batch64|serve]

A step matches a stream of distinct synthetic scenes of the workload (8 by default), each already
resident in HBM, against the whole template bank, with 4 images in flight on their own streams (one
captured graph per stream); `value` is milliseconds per image.  Each match covers the per-image
pipeline (gradients, spreading, linear memories for both pyramid levels), the coarse
gather-accumulate with threshold compaction, the refinement, and the return of the canonical match
list to the host; every streamed result is checked against the same image matched alone.  The
latency of one image matched alone (graph replay) is reported beside it.

Default workload: a 1024x1024 scene with 8 rotated/scaled copies of a polygon, a bank of 360
rotations x 256 level-0 features (128 at level 1), T = {4, 8}, threshold 80 -- the plain-GPU
drop-in configuration.  ``--workload bank36k`` runs the 2048x2048 scene against the 36,000-template
bank (20 classes x 360 rotations x 5 scales, features U[32, 1024] with 1% at 8191).

Prints ONE JSON line on stdout; progress goes to stderr.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

METRIC = "ms/image at 1024² × 360 templ and 36k-templ bank, batched vs plain; served p99 ms"
HBM_PEAK_GBS = 8000.0
# Per-clock rates of MI355X's 256 CUs at 2.4 GHz for the resources that bind the matching kernels:
# the vector L1 serves one access (one lane quad's cache line: TCP_TOTAL_CACHE_ACCESSES) and returns up to
# 128 bytes per CU per clock (614.4 G accesses/s, 78.6 TB/s), and the LDS serves 32 banks (one dword atomic
# each) per CU per clock (19.7 T atomics/s).
CUS, CLOCK_HZ = 256, 2.4e9
L1_ACCESS_PEAK_G = CUS * CLOCK_HZ / 1e9
L1_PEAK_GBS = CUS * 128 * CLOCK_HZ / 1e9
LDS_ATOMIC_PEAK_G = CUS * 32 * CLOCK_HZ / 1e9


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def bgr3(gray):
    return np.ascontiguousarray(np.stack([gray, gray, gray], -1))


def build_workload(name):
    from shape_based_matching_amd import synth
    if name == "copies1024":
        img, _, placed = synth.scene_copies(1)
        _, _, poly, _ = synth.scene_plain(0)
        bank = {"poly": [synth.bank_template(poly, float(a), 1.0, 60, 256) for a in range(360)]}
        cfg = dict(workload="1024x1024 scene with 8 rotated/scaled polygon copies (blur, 2% salt-and-pepper); "
                            "bank 360 rotations x 256 level-0 / 128 level-1 features placed on the rotated outline (synthetic, "
                            "not extracted from an image); T={4,8}; threshold 80",
                   image=[1024, 1024], templates=360, features_level0=256, T=[4, 8], threshold=80, seed=1)
        return img, bank, 80.0, cfg
    if name == "copies1024_16":
        img, _, placed = synth.scene_copies(1)
        _, _, poly, _ = synth.scene_plain(0)
        bank = {"poly": [synth.bank_template(poly, float(a), 1.0, 60, 256, bins=16) for a in range(360)]}
        cfg = dict(workload="16-orientation variant (line2Dup_16bit_ori.cpp): the 1024x1024 copies scene as BGR "
                            "(gray in all three channels); bank 360 rotations x 256 level-0 / 128 level-1 features "
                            "with 16-bin labels; T={4,8}; threshold 60",
                   image=[1024, 1024, 3], templates=360, features_level0=256, T=[4, 8], threshold=60, seed=1,
                   orientations=16)
        return bgr3(img), bank, 60.0, cfg
    if name == "bank36k":
        bank = {}
        polys = {}
        for cid, pyr, poly in synth.synthetic_bank(seed=2):
            bank.setdefault(cid, []).append(pyr)
            polys[cid] = poly
        img = synth.scene_with_bank_shapes(3, 2048, list(polys.values()))
        _POLYS["bank36k"] = list(polys.values())  # the stream's scenes use the same polygons
        n = sum(len(v) for v in bank.values())
        cfg = dict(workload="2048x2048 scene with 24 bank polygons; bank 20 classes x 360 rotations x 5 scales, "
                            "level-0 features U[32,1024] with 1% at 8191, level-1 ceil(n/2); T={4,8}; threshold 80",
                   image=[2048, 2048], templates=n, T=[4, 8], threshold=80, seed=2)
        return img, bank, 80.0, cfg
    if name == "batch64":
        imgs, poly = synth.mixed_batch(4, 64)
        bank = {"poly": [synth.bank_template(poly, a * 0.5, 1.0, 60, 512) for a in range(720)]}
        cfg = dict(workload="64 scenes of mixed sizes (640x480, 1280x960, 2592x1936) with 10-40 shapes and 3 copies of "
                            "the polygon; bank 720 rotations (0.5 deg) x 512 level-0 / 256 level-1 features; T={4,8}; "
                            "threshold 80; one captured graph per size bucket, 4 streams",
                   images=64, templates=720, features_level0=512, T=[4, 8], threshold=80, seed=4)
        return imgs, bank, 80.0, cfg
    if name == "serve":
        bank, polys = {}, []
        for cid, pyr, poly in synth.synthetic_bank(seed=5, n_classes=25, angles=np.arange(200) * 1.8, scales=(1.0,),
                                                   nmin=64, nmax=2048, big_frac=0.0):
            bank.setdefault(cid, []).append(pyr)
            if not polys or polys[-1] is not poly:
                polys.append(poly)
        g = synth.rng(6)
        pool = []
        for i in range(24):
            rows, cols = synth.BATCH_SIZES[i % 3]
            pool.append(synth.scene_sized(6000 + i, rows, cols, polys[int(g.integers(0, len(polys)))]))
        n = sum(len(v) for v in bank.values())
        cfg = dict(workload="Poisson stream over a pool of 24 mixed-size scenes (640x480, 1280x960, 2592x1936; 10-40 "
                            "shapes and 3 copies of a bank polygon each); resident bank 25 classes x 200 rotations "
                            "(1.8 deg) = 5000 templates, level-0 features U[64,2048], level-1 ceil(n/2); T={4,8}; "
                            "threshold 80; continuous batching over several slots under a fixed HBM budget",
                   templates=n, pool_images=len(pool), T=[4, 8], threshold=80, seed=5)
        return pool, bank, 80.0, cfg
    raise SystemExit(f"unknown workload {name}")


def measured_traffic(workload, kernel):
    """HBM bytes per launch from the committed rocprofv3 --pmc summaries (FETCH_SIZE + WRITE_SIZE,
    scaled by a known-size copy kernel), or None when no summary covers this kernel."""
    prof = os.path.join(HERE, "profiles")
    if not os.path.isdir(prof):
        return None
    for fn in sorted(os.listdir(prof)):
        if not fn.endswith(".md"):
            continue
        for line in open(os.path.join(prof, fn)):
            parts = line.split()
            if len(parts) == 4 and parts[0] == "TRAFFIC" and parts[1] == workload and parts[2] == kernel:
                return int(parts[3])
    return None


_POLYS = {}


def stream_scene(workload, k):
    """Scene k (k >= 1) of a stream for a single-image workload: same generator, another seed."""
    from shape_based_matching_amd import synth
    if workload == "copies1024":
        return synth.scene_copies(100 + k)[0]
    if workload == "copies1024_16":
        return bgr3(synth.scene_copies(100 + k)[0])
    if workload == "bank36k":
        if "bank36k" not in _POLYS:
            polys = {}
            for cid, _, poly in synth.synthetic_bank(seed=2):
                polys[cid] = poly
            _POLYS["bank36k"] = list(polys.values())
        return synth.scene_with_bank_shapes(100 + k, 2048, _POLYS["bank36k"])
    raise SystemExit(f"no stream for {workload}")


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_threads():
    """Host threads for the CPU baseline: the process's CPU share, at most 16 (one GPU's share of its host)."""
    return min(16, len(os.sched_getaffinity(0)))


def cpu_baseline_fast(images, bank, thr, T, seconds, min_runs=10, check=None):
    """The vectorised, OpenMP restatement of the reference's CPU path (oracle/sbm_cpu_fast.c) timed on this
    host: preprocessing on one thread as the reference's, templates in parallel over cpu_threads() (as
    line2Dup.cpp:1137-1141).  Two warm-up runs, then at least min_runs runs (cycling over `images`) within
    about `seconds`; medians of the total and of the reference's two stage timers.  check: (index, GPU list)
    compared with the baseline's list for that image."""
    import oracle
    threads = cpu_threads()
    fb = oracle.FlatBank(bank, len(T))
    log("cpu baseline (vectorised %s, %d threads)" % (oracle.fast_isa(), threads))
    for k in range(2):
        oracle.match_fast(images[k % len(images)], T, thr, fb, nthreads=threads)
    tot, pre, mat = [], [], []
    t_start = time.perf_counter()
    i = 0
    equal = None
    while len(tot) < min_runs or time.perf_counter() - t_start < seconds:
        img = images[i % len(images)]
        t0 = time.perf_counter()
        res, st = oracle.match_fast(img, T, thr, fb, nthreads=threads)
        tot.append(time.perf_counter() - t0)
        pre.append(st[0])
        mat.append(st[1])
        if check is not None and i % len(images) == check[0]:
            xyct, sim = check[1]
            equal = len(res) == len(sim) and all(
                (r[0], r[1], np.float32(r[2]), r[4]) == (int(a[0]), int(a[1]), s, int(a[3]))
                for r, a, s in zip(res, xyct, sim))
        i += 1
        if len(tot) >= 400 or (len(tot) >= min_runs and time.perf_counter() - t_start > seconds):
            break
    out = {"value": round(float(np.median(tot)) * 1e3, 3), "unit": "ms/image", "cores": threads, "kind": "port",
           "cpu_model": cpu_model(), "isa": oracle.fast_isa(),
           "response_maps_ms": round(float(np.median(pre)) * 1e3, 3),
           "templ_match_ms": round(float(np.median(mat)) * 1e3, 3),
           "sample": f"{len(tot)} runs over {min(len(images), len(tot))} image(s) of the workload after 2 warm-ups "
                     f"(median): oracle/sbm_cpu_fast.c, the reference's method vectorised ({oracle.fast_isa()}), "
                     f"response maps on one thread and templates over {threads} threads as "
                     f"line2Dup.cpp:1061-1091, 1137-1141"}
    if equal is not None:
        out["matches_equal_gpu"] = bool(equal)
    return out


TUNING_PARAMS = ("list_threads", "list_segment", "list_blocks_per_cu", "list_order", "list_split", "refine_waves", "refine_group",
                 "refine_blocks_per_cu", "refine_layout")


def set_params(det, spec):
    """--set name=value,... through sbm_set_tuning_param (list_blocks_per_cu first to 1, so that any order of a
    valid set passes the per-CU limits on the way)."""
    if not spec:
        return
    pairs = [kv.split("=") for kv in spec.split(",") if kv]
    det.set_tuning_param("list_blocks_per_cu", 1)
    for k, v in sorted(pairs, key=lambda kv: kv[0] == "list_blocks_per_cu"):
        det.set_tuning_param(k.strip(), int(v))


def kernel_work(name, rows, cols, ch, gather_bytes, refine_feats, list_entries, nib_level=-1, planes=8, levels=2,
                nimg=1):
    """Algorithmic work of one launch of a named kernel in the unit of the resource that binds it:
    (amount, unit, peak, bound, what).  Preprocessing kernels stream from and to HBM (input + output bytes);
    the gather coarse kernel reads its gather bytes from the vector L1 (half a byte per position on the
    nibble-packed level); the nonzero-response coarse kernel makes one LDS atomic per list entry it reads;
    the refinement reads one L1 line per window row: 16 per feature per candidate.
    nib_level: the level whose linear memories are also written two positions per byte; planes: orientations;
    levels: pyramid levels (the gradient kernel of every level but the last also writes the next level's image);
    nimg: images of the launch (gather_bytes is one image's; refine_feats and list_entries count the launch)."""
    lvl = int(name.rsplit("_l", 1)[1]) if "_l" in name and name.rsplit("_l", 1)[1].isdigit() else 0
    px = (rows >> lvl) * (cols >> lvl) * nimg
    gather_bytes = gather_bytes * nimg
    qb = planes // 8
    hbm = (HBM_PEAK_GBS, "hbm", "GB/s")
    if name.startswith("gradient"):
        if lvl < levels - 1:
            return (px * ch + px * qb + px // 4 * ch, hbm[2], hbm[0], hbm[1],
                    "HBM bytes: image in, quantized orientations and the next level's image out")
        return px * ch + px * qb, hbm[2], hbm[0], hbm[1], "HBM bytes: image in, quantized orientations out"
    if name.startswith("spread_linearize"):
        return (px * qb + planes * px + (4 * px if lvl == nib_level and planes == 8 else 0), hbm[2], hbm[0], hbm[1],
                "HBM bytes: quantized in, linear memories (and the nibble copy) out")
    if name.startswith("pack_nibbles"):
        return planes * px + planes * px // 2, hbm[2], hbm[0], hbm[1], "HBM bytes: linear memories in, nibbles out"
    if name.startswith("pyr_down"):
        return 4 * px * ch + px * ch, hbm[2], hbm[0], hbm[1], "HBM bytes: level image in, half-size image out"
    if name.startswith("coarse") and list_entries:
        return (list_entries / 1e9, "G atomics/s", LDS_ATOMIC_PEAK_G, "lds_atomics",
                "one LDS dword atomic per nonzero-response list entry read")
    if name.startswith("coarse"):
        per = 0.5 if nib_level >= 0 else 1.0
        return (gather_bytes * per, "GB/s", L1_PEAK_GBS, "l1",
                "vector-L1 bytes: one response per feature per template position (a nibble on the packed level)")
    if name.startswith("refine") and lvl == 0:
        return (16 * refine_feats / 1e9, "G accesses/s", L1_ACCESS_PEAK_G, "l1_accesses",
                "vector-L1 accesses: 16 per feature per candidate (a 16 x 16 window is 64 lanes, 16 lane quads, each "
                "quad one access per cache line it touches; measured 16-19 per feature in every refinement layout)")
    return 0, hbm[2], hbm[0], hbm[1], "not counted"


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default=None,
                    choices=["copies1024", "bank36k", "batch64", "serve", "copies1024_16"],
                    help="one workload only (default: the 1024^2 headline, then the 36k bank and serving parts)")
    ap.add_argument("--steps-36k", type=int, default=10, help="timed steps of the 36k-bank part of the default run")
    ap.add_argument("--plain-steps-36k", type=int, default=2, help="plain-path steps of the 36k-bank part")
    ap.add_argument("--serve-loads", default="0.5,0.75", help="offered loads of the serving part of the default run")
    ap.add_argument("--slots", type=int, default=4, help="streams in flight for --workload batch64 / serve")
    ap.add_argument("--budget-mib", type=int, default=320,
                    help="device-memory budget of the server: the resident bank (about 64 MB for the serving bank) "
                         "and the slots")
    ap.add_argument("--loads", default="0.25,0.5,0.75,0.9",
                    help="offered request rates for --workload serve, as fractions of the measured saturation rate")
    ap.add_argument("--serve-seconds", type=float, default=4.0, help="arrival window per offered rate")
    ap.add_argument("--cpu-seconds", type=float, default=12.0, help="bounded CPU-baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--plain-steps", type=int, default=10)
    ap.add_argument("--images-per-launch", type=int, default=8,
                    help="images of one size per launch in the stream (0: half of the stream, two launches in turn)")
    ap.add_argument("--stream-slots", type=int, default=2, help="launches in flight in the stream")
    ap.add_argument("--stream-images", type=int, default=48,
                    help="distinct scenes per step of the throughput measurement (0: single-image steps only)")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--coarse-mode", type=int, default=1,
                    help="batched coarse stage: 0 gather kernel, 1 auto (default), 2 nonzero-response list when it fits")
    ap.add_argument("--tuning", default=None, help="fixed coarse tile shape 'ppl,wpb,bpc[,gather[,split]]' (skips the autotuner)")
    ap.add_argument("--set", default=None,
                    help="named launch-shape parameters 'name=value,...' (sbm_set_tuning_param; with --tuning or "
                         "--no-autotune they stay as given)")
    ap.add_argument("--only-stream", action="store_true",
                    help="run only the warm-up and the timed stream steps (for profiling the multi-image launches)")
    ap.add_argument("--copy-probe", type=int, default=0,
                    help="run the calibration copy kernel on this many MiB first (for rocprofv3 --pmc scaling)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU: the matching path is single-GPU in this build (template or image "
              "sharding across GPUs is a design note, not implemented)", file=sys.stderr)
        sys.exit(2)

    if args.workload is not None:
        print(json.dumps(run_workload(args, args.workload)), flush=True)
        return
    # the whole metric: the 1024^2 x 360 headline (value), then the 36k-template bank (batched vs plain)
    # and the served p50/p99 over the 5,000-template bank, as keys of the same line
    out = run_workload(args, "copies1024")
    sub = argparse.Namespace(**vars(args))
    sub.steps, sub.warmup, sub.stream_images, sub.plain_steps = args.steps_36k, 2, 32, args.plain_steps_36k
    b36 = run_workload(sub, "bank36k")
    out["bank36k"] = {k: b36[k] for k in ("value", "unit", "config", "images_per_s", "latency_ms_per_image",
                                          "graph_device_ms_per_image", "plain_ms_per_image",
                                          "batched_speedup_vs_plain", "matches", "coarse_candidates", "coarse_path",
                                          "kernels_ms", "roofline", "single_image_roofline", "stream_launch",
                                          "cpu_baseline")}
    out["bank36k"]["batched_ms_per_image"] = b36["value"]
    out["bank36k"]["stream_speedup_vs_plain"] = (None if b36["plain_ms_per_image"] is None
                                                 else round(b36["plain_ms_per_image"] / b36["value"], 2))
    sub = argparse.Namespace(**vars(args))
    sub.loads, sub.serve_seconds = args.serve_loads, args.serve_seconds
    srv = run_workload(sub, "serve")
    out["serve"] = {k: srv[k] for k in ("value", "unit", "config", "saturation_images_per_s", "sweep",
                                        "plain_ms_per_image", "served_equal_plain", "cpu_baseline")}
    out["serve"]["p99_ms"] = srv["value"]
    out["serve"]["p50_ms"] = srv["ms_per_step"]
    print(json.dumps(out), flush=True)


def run_workload(args, workload):
    import shape_based_matching_amd as sbm

    log("building workload", workload)
    img, bank, thr, cfg = build_workload(workload)
    if workload == "batch64":
        return bench_batch(args, img, bank, thr, cfg)
    if workload == "serve":
        return bench_serve(args, img, bank, thr, cfg)
    rows, cols = img.shape[:2]
    ch = 1 if img.ndim == 2 else img.shape[2]
    planes = cfg.get("orientations", 8)
    det = sbm.Detector16(256, [4, 8]) if planes == 16 else sbm.Detector(256, [4, 8])
    for cid in sorted(bank):
        for pyr in bank[cid]:
            det.add_template_pyramid(cid, pyr)
    log("bank", det.numTemplates(), "templates in", det.numClasses(), "classes")
    det.set_coarse_mode(args.coarse_mode)
    det.set_images_per_launch(args.images_per_launch)
    dimg = det.upload_image(img)

    # warm-up (also uploads the bank to HBM), then the tile-shape autotuner (untimed)
    tuning = None
    if args.copy_probe:
        det.copy_probe(args.copy_probe << 20)
    if args.tuning:
        vals = [int(v) for v in args.tuning.split(",")]
        det.set_tuning(*vals)
        tuning = dict(zip(("positions_per_lane", "waves_per_block", "blocks_per_cu", "gather", "split"), det.get_tuning()))
    set_params(det, args.set)
    if args.set:
        tuning = dict(tuning or {}, **{k: det.get_tuning_param(k) for k in TUNING_PARAMS})
    for _ in range(max(args.warmup, 1)):
        det.match_image_arrays(dimg, thr, mode=sbm.MODE_BATCHED)
    # the stream's distinct scenes, resident in HBM (the autotuner's multi-image launches take the first of them)
    scenes = [img] + [stream_scene(workload, k) for k in range(1, args.stream_images)]
    dimgs = [dimg] + [det.upload_image(im) for im in scenes[1:]]
    if not args.tuning and not args.no_autotune:
        log("autotune")
        best, _ = det.autotune(dimg, thr, iters=2, batch=dimgs[1:args.images_per_launch])
        tuning = dict(zip(("positions_per_lane", "waves_per_block", "blocks_per_cu", "gather", "split"), best))
        tuning.update({k: det.get_tuning_param(k) for k in TUNING_PARAMS})
        tuning["coarse_mode"] = det.get_coarse_mode()
        log("tuning", tuning)
        for _ in range(2):
            det.match_image_arrays(dimg, thr, mode=sbm.MODE_BATCHED)
    det.synchronize()
    if args.only_stream:
        for _ in range(2):
            det.match_images_arrays(dimgs, thr, slots=args.stream_slots)
        det.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            det.match_images_arrays(dimgs, thr, slots=args.stream_slots)
        det.synchronize()
        step_ms = (time.perf_counter() - t0) * 1e3 / args.steps
        return {"metric": METRIC, "value": round(step_ms / len(dimgs), 4), "unit": "ms/image", "n_gpus": 1,
                "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(step_ms, 4),
                "higher_is_better": False, "config": dict(cfg, tuning=tuning, images_per_step=len(dimgs),
                                                          images_per_launch=args.images_per_launch,
                                                          slots=args.stream_slots, only_stream=True)}
    log("timing", args.steps, "single-image steps")
    t0 = time.perf_counter()
    for _ in range(args.steps):
        xyct, sim, _ = det.match_image_arrays(dimg, thr, mode=sbm.MODE_BATCHED)
    det.synchronize()
    t1 = time.perf_counter()
    ms_single = (t1 - t0) * 1e3 / args.steps
    n_matches = len(sim)

    # throughput: a stream of distinct scenes of the same workload, `slots` in flight on their own
    # streams (one captured graph each); every result is checked against the image matched alone
    stream = None
    if args.stream_images > 0:
        alone = [det.match_image_arrays(d, thr, mode=sbm.MODE_BATCHED) for d in dimgs]
        for _ in range(2):
            res = det.match_images_arrays(dimgs, thr, slots=args.stream_slots)
        for k, (a, b) in enumerate(zip(alone, res)):
            if not _same(a, b):
                raise SystemExit(f"stream image {k}: batched result differs from the image matched alone")
        det.synchronize()
        log("timing", args.steps, "stream steps of", len(dimgs), "images")
        t0 = time.perf_counter()
        for _ in range(args.steps):
            det.match_images_arrays(dimgs, thr, slots=args.stream_slots)
        det.synchronize()
        step_ms = (time.perf_counter() - t0) * 1e3 / args.steps
        stream = {"images_per_step": len(dimgs), "slots": args.stream_slots,
                  "images_per_launch": args.images_per_launch, "ms_per_step": round(step_ms, 4),
                  "ms_per_image": round(step_ms / len(dimgs), 4),
                  "images_per_s": round(1e3 * len(dimgs) / step_ms, 1)}
    ms_batched = stream["ms_per_image"] if stream else ms_single

    # the same stream from host memory: each launch's images staged through pinned memory, one DMA per
    # launch, the copies of one launch overlapping the other launch's kernels (a reference caller hands
    # over host images, line2Dup.h:154-156); reported beside value, never as value
    host_stream = None
    if stream:
        hres = det.match_host_images_arrays(scenes, thr, slots=args.stream_slots)
        for k, (a, b) in enumerate(zip(alone, hres)):
            if not _same(a, b):
                raise SystemExit(f"host stream image {k}: result differs from the image matched alone")
        log("timing", args.steps, "host-input stream steps")
        t0 = time.perf_counter()
        for _ in range(args.steps):
            det.match_host_images_arrays(scenes, thr, slots=args.stream_slots)
        hstep = (time.perf_counter() - t0) * 1e3 / args.steps
        host_stream = {"ms_per_step": round(hstep, 4), "ms_per_image": round(hstep / len(scenes), 4),
                       "images_per_s": round(1e3 * len(scenes) / hstep, 1)}

    # plain path (one coarse launch per template)
    plain_ms = None
    if args.plain_steps > 0:
        det.match_image_arrays(dimg, thr, mode=sbm.MODE_PLAIN)
        det.synchronize()
        log("timing", args.plain_steps, "plain steps")
        t0 = time.perf_counter()
        for _ in range(args.plain_steps):
            pxy, psim, _ = det.match_image_arrays(dimg, thr, mode=sbm.MODE_PLAIN)
        det.synchronize()
        plain_ms = (time.perf_counter() - t0) * 1e3 / args.plain_steps
        if not (np.array_equal(pxy, xyct) and np.array_equal(psim.view(np.uint32), sim.view(np.uint32))):
            raise SystemExit("plain and batched match lists differ")

    # host-buffer input (PCIe-inclusive), for reference only
    t0 = time.perf_counter()
    nh = max(3, args.steps // 5)
    for _ in range(nh):
        det.match_arrays(img, thr, mode=sbm.MODE_BATCHED)
    pcie_ms = (time.perf_counter() - t0) * 1e3 / nh

    # device span of one replayed graph (input copy to result copy), HIP events on the detector's stream
    det.set_timing(2)
    spans = []
    for _ in range(max(5, args.steps // 2)):
        det.match_image_arrays(dimg, thr, mode=sbm.MODE_BATCHED)
        spans.append(det.last_device_ms())
    graph_ms = float(np.median(spans))

    # per-kernel device times with HIP events on the detector's stream
    det.set_timing(True)
    acc = {}
    nk = max(5, args.steps // 2)
    cand = []
    for _ in range(nk):
        det.match_image_arrays(dimg, thr, mode=sbm.MODE_BATCHED)
        for name, ms in det.last_kernel_times():
            acc.setdefault(name, []).append(ms)
        cand.append(det.last_timing()["coarse_candidates"])
    path = det.last_coarse_path()
    entries = det.last_coarse_entries()
    kernels = {k: float(np.mean(v)) for k, v in acc.items()}
    dominant = max(kernels, key=kernels.get)
    refine_feats = det.last_refine_features()
    det.set_timing(False)
    gather = det.coarse_gather_bytes(rows, cols)
    nib_level = det.pyramidLevels() - 1 if det.get_tuning()[3] == 3 else -1
    list_entries = entries if path[0] else 0
    per_kernel = {}
    for k, ms in sorted(kernels.items()):
        w, unit, peak, bound, what = kernel_work(k, rows, cols, ch, gather, refine_feats, list_entries, nib_level,
                                                 planes, det.pyramidLevels())
        rate = w / (ms * 1e-3) / (1e9 if unit == "GB/s" else 1.0)
        per_kernel[k] = {"ms": round(ms, 5), "bound": bound, "achieved": round(rate, 2), "unit": unit,
                         "peak": round(peak, 1), "frac": round(rate / peak, 4) if w else None, "work": what}
    w, unit, peak, bound, what = kernel_work(dominant, rows, cols, ch, gather, refine_feats, list_entries, nib_level,
                                             planes, det.pyramidLevels())
    dur_s = kernels[dominant] * 1e-3
    achieved = w / dur_s / (1e9 if unit == "GB/s" else 1.0)
    traffic = measured_traffic(workload, dominant)
    roofline = {"kernel": dominant, "bound": bound, "achieved": round(achieved, 2), "peak": round(peak, 1),
                "unit": unit, "frac": round(achieved / peak, 4), "traffic": traffic,
                "avg_launch_ms": round(kernels[dominant], 5), "work": what,
                "peak_basis": ("MI355X HBM3E 8 TB/s" if bound == "hbm" else
                               "256 CUs x 2.4 GHz x per-clock rate (L1: one access and up to 128 bytes per CU "
                               "per clock; LDS: 32 dword banks per CU per clock)")}
    if traffic:
        roofline["hbm_GBs"] = round(traffic / dur_s / 1e9, 1)
        roofline["hbm_frac"] = round(traffic / dur_s / 1e9 / HBM_PEAK_GBS, 4)
    if dominant.startswith("coarse") or dominant.startswith("refine"):
        # the gather-equivalent rate: one uint8 read + add per feature per position, whatever is really read
        gb = gather if dominant.startswith("coarse") else 256 * refine_feats
        roofline["gather_equivalent_GBs"] = round(gb / dur_s / 1e9, 1)
    if dominant.startswith("coarse") and path[0]:
        roofline["list_entries_read"] = int(entries)
        roofline["list_GBs"] = round(4 * entries / dur_s / 1e9, 1)
    coarse_name = [k for k in kernels if k.startswith("coarse")]
    coarse_rate = None
    if coarse_name:
        cs = kernels[coarse_name[0]] * 1e-3
        coarse_rate = {"gather_bytes": int(gather), "avg_launch_ms": round(kernels[coarse_name[0]], 5),
                       "gather_equivalent_GBs": round(gather / cs / 1e9, 1)}

    # one launch of the stream (images_per_launch distinct scenes, every kernel covering all of them), kernel by
    # kernel with HIP events on its pipeline's stream (the launch runs eagerly); its dominant kernel's roofline
    stream_launch = None
    if stream:
        nl = min(args.images_per_launch, len(dimgs))
        det.set_timing(True)
        acc8, ent8, feat8 = {}, [], []
        for _ in range(max(5, args.steps // 2)):
            det.match_images_arrays(dimgs[:nl], thr, slots=1)
            for name, ms in det.last_kernel_times():
                acc8.setdefault(name, []).append(ms)
            ent8.append(det.last_coarse_entries())
            feat8.append(det.last_refine_features())
        path8 = det.last_coarse_path()
        det.set_timing(False)
        k8 = {k: float(np.mean(v)) for k, v in acc8.items()}
        dom8 = max(k8, key=k8.get)
        e8 = int(np.mean(ent8)) if path8[0] else 0
        f8 = int(np.mean(feat8))
        w, unit, peak, bound, what = kernel_work(dom8, rows, cols, ch, gather, f8, e8, nib_level, planes,
                                                 det.pyramidLevels(), nimg=nl)
        d8 = k8[dom8] * 1e-3
        a8 = w / d8 / (1e9 if unit == "GB/s" else 1.0)
        rl8 = {"kernel": dom8, "bound": bound, "achieved": round(a8, 2), "peak": round(peak, 1), "unit": unit,
               "frac": round(a8 / peak, 4), "traffic": measured_traffic(workload + "_x%d" % nl, dom8),
               "avg_launch_ms": round(k8[dom8], 5), "work": what,
               "launch": f"one launch of the stream behind value: {nl} distinct scenes, every kernel covering all of them",
               "peak_basis": roofline["peak_basis"] if bound == roofline["bound"] else
               ("MI355X HBM3E 8 TB/s" if bound == "hbm" else
                "256 CUs x 2.4 GHz x per-clock rate (L1: one access and up to 128 bytes per CU per clock; LDS: "
                "32 dword banks per CU per clock)")}
        if rl8["traffic"]:
            rl8["hbm_GBs"] = round(rl8["traffic"] / d8 / 1e9, 1)
            rl8["hbm_frac"] = round(rl8["traffic"] / d8 / 1e9 / HBM_PEAK_GBS, 4)
        if dom8.startswith("refine"):
            rl8["gather_equivalent_GBs"] = round(256 * f8 / d8 / 1e9, 1)
        if dom8.startswith("coarse") and path8[0]:
            rl8["list_entries_read"] = e8
        stream_launch = {"images": nl, "coarse_path": "nonzero_list" if path8[0] else "gather",
                         "kernels_ms": {k: round(v, 5) for k, v in sorted(k8.items())},
                         "device_ms": round(sum(k8.values()), 5), "roofline": rl8}

    cpu = None
    if not args.no_cpu and planes == 8:
        cpu = cpu_baseline_fast([img], bank, thr, [4, 8], args.cpu_seconds, check=(0, (xyct, sim)))
    elif not args.no_cpu:  # the 16-orientation variant: its scalar oracle, templates over the host threads
        import oracle
        cores = cpu_threads()
        log("cpu baseline (16-orientation oracle, %d threads), ~%.0f s" % (cores, args.cpu_seconds))
        times = []
        t_start = time.perf_counter()
        while True:
            t0 = time.perf_counter()
            ref = oracle.match16(img, [4, 8], thr, bank, nthreads=cores)
            times.append(time.perf_counter() - t0)
            if time.perf_counter() - t_start > args.cpu_seconds or len(times) >= 50:
                break
        ok = len(ref) == n_matches and all(
            (r[0], r[1], np.float32(r[2]), r[4]) == (int(a[0]), int(a[1]), s, int(a[3]))
            for r, a, s in zip(ref, xyct, sim))
        cpu = {"value": round(float(np.median(times)) * 1e3, 3), "unit": "ms/image", "cores": cores, "kind": "port",
               "cpu_model": cpu_model(),
               "sample": f"{len(times)} full images of the same workload (median), the scalar 16-orientation "
                         f"oracle: single-threaded preprocessing, OpenMP over templates",
               "matches_equal_gpu": bool(ok)}

    out = {
        "metric": METRIC,
        "value": round(ms_batched, 4),
        "unit": "ms/image",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(stream["ms_per_step"] if stream else ms_single, 4),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u8",
        "data": "synthetic (seeded numpy PCG64 scenes and banks)",
        "config": dict(cfg, mode="batched", tuning=tuning,
                       step=(f"{stream['images_per_step']} distinct scenes of this workload resident in HBM, "
                             f"{stream['images_per_launch']} per launch (every kernel covers all of a launch's "
                             f"images), {stream['slots']} launches in flight; value = ms per image") if stream else
                       "one image; value = ms per image"),
        "images_per_s": round(1e3 / ms_batched, 1),
        "latency_ms_per_image": round(ms_single, 4),
        "stream": stream,
        "plain_ms_per_image": None if plain_ms is None else round(plain_ms, 4),
        "batched_speedup_vs_plain": None if plain_ms is None else round(plain_ms / ms_single, 2),
        "graph_device_ms_per_image": round(graph_ms, 4),
        "pcie_inclusive_ms_per_image": round(pcie_ms, 4),
        "host_input_stream": host_stream,
        "matches": n_matches,
        "coarse_candidates": int(np.mean(cand)),
        "coarse_path": {"nonzero_list": bool(path[0]), "nonzero_responses": int(path[1]), "list_entries_read": int(entries)},
        "kernels_ms": {k: round(v, 5) for k, v in sorted(kernels.items())},
        "kernel_roofline": per_kernel,
        "coarse_gather": coarse_rate,
        # the dominant kernel of the launches value is measured on (the stream's), else of the single-image launch
        "roofline": stream_launch["roofline"] if stream_launch else roofline,
        "single_image_roofline": roofline,
        "stream_launch": stream_launch,
        "cpu_baseline": cpu,
    }
    return out


def bench_batch(args, imgs, bank, thr, cfg):
    """Images/s of a mixed-size batch: captured graphs per size bucket on several streams, against the
    same images matched one at a time (eager launches) and the CPU oracle."""
    import shape_based_matching_amd as sbm
    det = sbm.Detector(512, [4, 8])
    for cid in sorted(bank):
        for pyr in bank[cid]:
            det.add_template_pyramid(cid, pyr)
    if args.tuning:
        det.set_tuning(*(int(v) for v in args.tuning.split(",")))
    set_params(det, args.set)
    dimgs = [det.upload_image(im) for im in imgs]
    for _ in range(max(args.warmup, 1)):
        det.match_images_arrays(dimgs, thr, slots=args.slots)
    steps = max(1, args.steps // 10)
    log("timing", steps, "batches of", len(imgs))
    t0 = time.perf_counter()
    for _ in range(steps):
        res = det.match_images_arrays(dimgs, thr, slots=args.slots)
    ms_batch = (time.perf_counter() - t0) * 1e3 / (steps * len(imgs))
    t0 = time.perf_counter()
    single = [det.match_image_arrays(d, thr, mode=sbm.MODE_BATCHED) for d in dimgs]
    ms_single = (time.perf_counter() - t0) * 1e3 / len(imgs)
    same = all(np.array_equal(a[0], b[0]) and np.array_equal(a[1].view(np.uint32), b[1].view(np.uint32))
               for a, b in zip(res, single))
    if not same:
        raise SystemExit("batched-graph and single-image match lists differ")
    plain_ms = None
    if args.plain_steps > 0:
        t0 = time.perf_counter()
        for d in dimgs[:8]:
            det.match_image_arrays(d, thr, mode=sbm.MODE_PLAIN)
        plain_ms = (time.perf_counter() - t0) * 1e3 / 8
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline_fast(imgs, bank, thr, [4, 8], args.cpu_seconds, min_runs=len(imgs))

    out = {"metric": METRIC, "value": round(ms_batch, 4), "unit": "ms/image", "n_gpus": 1, "steps": steps,
           "warmup": args.warmup, "ms_per_step": round(ms_batch * len(imgs), 3), "higher_is_better": False,
           "scaling": "weak", "vs_baseline": None, "dtype": "u8", "data": "synthetic (seeded numpy PCG64 scenes and banks)",
           "config": dict(cfg, slots=args.slots), "images_per_s": round(1e3 / ms_batch, 1),
           "single_image_ms": round(ms_single, 4), "plain_ms_per_image": None if plain_ms is None else round(plain_ms, 3),
           "matches": int(sum(len(r[1]) for r in res)), "cpu_baseline": cpu}
    return out


def _same(a, b):
    return np.array_equal(a[0], b[0]) and np.array_equal(a[1].view(np.uint32), b[1].view(np.uint32))


def bench_serve(args, pool, bank, thr, cfg):
    """Continuous batching under Poisson arrivals: p50/p99 latency (submission to completion, host
    image in, match list out) and achieved images/s at several offered rates, every result checked
    bit-exact against the plain path for the same image."""
    import threading

    import shape_based_matching_amd as sbm
    det = sbm.Detector(2048, [4, 8])
    for cid in sorted(bank):
        for pyr in bank[cid]:
            det.add_template_pyramid(cid, pyr)
    log("bank", det.numTemplates(), "templates")
    dimgs = [det.upload_image(im) for im in pool]
    if args.tuning:
        det.set_tuning(*(int(v) for v in args.tuning.split(",")))
    set_params(det, args.set)
    if not args.tuning and not args.no_autotune:
        big = max(range(len(pool)), key=lambda i: pool[i].size)
        det.match_image_arrays(dimgs[big], thr)
        best, _ = det.autotune(dimgs[big], thr, iters=1)
        log("tuning", best)
    # the reference lists: the plain path (one coarse launch per template); with --plain-steps 0 (profiling
    # runs) the batched path one image at a time
    mode = sbm.MODE_PLAIN if args.plain_steps > 0 else sbm.MODE_BATCHED
    log("plain path for every pool image" if args.plain_steps > 0 else "batched single-image path for every pool image")
    t0 = time.perf_counter()
    plain = [det.match_image_arrays(d, thr, mode=mode) for d in dimgs]
    plain_ms = (time.perf_counter() - t0) * 1e3 / len(dimgs) if args.plain_steps > 0 else None
    for d in dimgs:
        d.close()
    budget = args.budget_mib << 20
    srv = sbm.Server(det, slots=args.slots, hbm_budget_bytes=budget)
    mismatches = [0]

    def closed_loop(n):
        t0 = time.perf_counter()
        ids = [(srv.submit(pool[k % len(pool)], thr), k % len(pool)) for k in range(n)]
        lat = []
        for rid, k in ids:
            res, ms = srv.wait_arrays(rid)
            mismatches[0] += not _same(res, plain[k])
            lat.append(ms)
        return n / (time.perf_counter() - t0), lat

    closed_loop(2 * len(pool))  # warm-up: one graph per size bucket and slot
    capacity, _ = closed_loop(4 * len(pool))
    log("saturation rate %.1f images/s" % capacity)
    g = np.random.default_rng(7)
    sweep = []
    for frac in (float(v) for v in args.loads.split(",")):
        rate = frac * capacity
        n = max(48, int(rate * args.serve_seconds))
        arrivals = np.cumsum(g.exponential(1.0 / rate, n))
        order = g.integers(0, len(pool), n)
        ids = []
        lat = [None] * n
        done = threading.Event()

        def waiter():
            j = 0
            while j < n:
                while j >= len(ids):
                    time.sleep(0.0002)
                rid, k, idx = ids[j]
                res, ms = srv.wait_arrays(rid)
                mismatches[0] += not _same(res, plain[k])
                lat[idx] = ms
                j += 1
            done.set()

        th = threading.Thread(target=waiter)
        th.start()
        t0 = time.perf_counter()
        for i in range(n):
            dt = t0 + arrivals[i] - time.perf_counter()
            if dt > 0:
                time.sleep(dt)
            ids.append((srv.submit(pool[order[i]], thr), int(order[i]), i))
        th.join()
        wall = time.perf_counter() - t0
        lat = np.array(lat)
        row = {"offered_fraction": frac, "offered_per_s": round(rate, 1), "requests": n,
               "achieved_per_s": round(n / wall, 1), "p50_ms": round(float(np.percentile(lat, 50)), 3),
               "p99_ms": round(float(np.percentile(lat, 99)), 3), "max_ms": round(float(lat.max()), 3)}
        log("load", row)
        sweep.append(row)
    st = srv.stats()
    srv.close()
    if mismatches[0]:
        raise SystemExit(f"{mismatches[0]} served results differ from the plain path")
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline_fast(pool, bank, thr, [4, 8], args.cpu_seconds, min_runs=len(pool))

    key = min(sweep, key=lambda r: abs(r["offered_fraction"] - 0.75))
    out = {"metric": METRIC, "value": key["p99_ms"], "unit": "ms (p99 latency at %.0f%% of the saturation rate)"
           % (100 * key["offered_fraction"]), "n_gpus": 1, "steps": int(sum(r["requests"] for r in sweep)),
           "warmup": 2 * len(pool), "ms_per_step": key["p50_ms"], "higher_is_better": False, "scaling": "weak",
           "vs_baseline": None, "dtype": "u8", "data": "synthetic (seeded numpy PCG64 scenes and banks)",
           "config": dict(cfg, slots=args.slots, budget_mib=args.budget_mib), "saturation_images_per_s": round(capacity, 1),
           "sweep": sweep, "plain_ms_per_image": None if plain_ms is None else round(plain_ms, 3),
           "served_equal_plain": True if plain_ms is not None else None,
           "server_stats": st, "cpu_baseline": cpu}
    return out


if __name__ == "__main__":
    main()
