"""Parity at the full workloads of the benchmark configurations.

* 2048x2048 scene against the 36,000-template bank (20 classes x 360 rotations x 5 scales, level-0
  features U[32, 1024] with 1% at 8191): the automatic coarse mode takes the nonzero-response list, and
  its match list equals the oracle's (line2Dup.cpp:1131-1268); the gather kernel, the list forced on, the
  plain per-template path and graph replays all return the same list.
* 8191 features at the coarse level itself, the 8-orientation limit (line2Dup.cpp:1166), on both coarse
  paths against the oracle, with sums close to 4 * 8191.
* 64 mixed-size scenes (640x480, 1280x960, 2592x1936) against 720 templates x 512 features: the batched
  multi-image call equals every scene matched alone, and every scene equals the oracle.
* Serving over the 5,000-template bank with features U[64, 2048]: every request equals the scene matched
  alone, whatever else is in flight, and a sample of two scenes per size equals the oracle.
"""
import os
import threading

import numpy as np
import pytest

import oracle
import shape_based_matching_amd as sbm
from helpers import oracle_tuples, tuples
from shape_based_matching_amd import synth

pytestmark = pytest.mark.gpu
NTHREADS = max(1, min(16, os.cpu_count() or 1))


def say(*a):
    print(*a, flush=True)


def detector_for(bank, nf=256):
    det = sbm.Detector(nf, [4, 8])
    for cid in sorted(bank):
        for pyr in bank[cid]:
            det.add_template_pyramid(cid, pyr)
    return det


# ------------------------------------------------------------------------------------------------
# 2048x2048 x 36,000 templates
# ------------------------------------------------------------------------------------------------
@pytest.fixture(scope="module")
def bank36k():
    bank, polys = {}, {}
    for cid, pyr, poly in synth.synthetic_bank(seed=2):
        bank.setdefault(cid, []).append(pyr)
        polys[cid] = poly
    img = synth.scene_with_bank_shapes(3, 2048, list(polys.values()))
    det = detector_for(bank)
    n = sum(len(v) for v in bank.values())
    big = sum(1 for v in bank.values() for p in v if len(p[0]["features"]) == 8191)
    say("36k bank:", n, "templates,", big, "with 8191 level-0 features")
    assert n == 36000 and big > 0
    say("oracle on the 2048x2048 scene,", NTHREADS, "threads")
    ref = oracle_tuples(oracle.match(img, [4, 8], 80, bank, nthreads=NTHREADS))
    say("oracle:", len(ref), "matches")
    assert len(ref) > 0
    return det, bank, img, ref


def test_bank36k_auto_mode_equals_oracle(bank36k):
    det, _, img, ref = bank36k
    det.set_coarse_mode(1)
    got = tuples(det.match(img, 80))
    path = det.last_coarse_path()
    say("auto mode: path", path, "matches", len(got))
    assert path[0], "the automatic mode takes the nonzero-response list on this workload"
    assert got == ref


def test_bank36k_every_coarse_path_and_graph_replays(bank36k):
    det, _, img, ref = bank36k
    dimg = det.upload_image(img)
    for mode in (0, 2, 1):
        det.set_coarse_mode(mode)
        for rep in range(3):  # eager, captured, replayed
            got = tuples(det.match_image(dimg, 80))
            assert got == ref, f"mode {mode} call {rep}"
        say("mode", mode, "path", det.last_coarse_path(), "ok")
    # the scatter's items in fixed XCD shares instead of from the work counter
    det.set_coarse_mode(2)
    det.set_tuning_param("list_order", 1)
    for rep in range(2):
        assert tuples(det.match_image(dimg, 80)) == ref, f"fixed shares, call {rep}"
    det.set_tuning_param("list_order", 0)
    det.set_coarse_mode(1)
    dimg.close()


def test_bank36k_every_refinement_layout(bank36k):
    """Every refinement layout on the full workload (the 8191-feature templates' candidates among them), single
    calls replayed and an eight-image launch of the scene, against the oracle."""
    det, _, img, ref = bank36k
    dimg = det.upload_image(img)
    for layout in (1, 2, 3, 4, 0):
        det.set_tuning_param("refine_layout", layout)
        for rep in range(2):  # captured, replayed
            assert tuples(det.match_image(dimg, 80)) == ref, f"layout {layout} call {rep}"
        det.set_images_per_launch(8)
        got = det.match_images([dimg] * 8, 80, slots=1)
        det.set_images_per_launch(0)
        assert all(tuples(x) == ref for x in got), f"layout {layout}, eight per launch"
        say("layout", layout, "ok")
    dimg.close()


def test_bank36k_plain_path_equals_oracle(bank36k):
    det, _, img, ref = bank36k
    got = tuples(det.match(img, 80, mode=sbm.MODE_PLAIN))
    assert got == ref


def test_bank36k_lower_threshold_modes_agree(bank36k):
    """Threshold 65 passes many more coarse candidates (the 8191-feature templates among them); the two
    coarse paths agree and the oracle agrees on the list."""
    det, bank, img, _ = bank36k
    out = {}
    for mode in (0, 2):
        det.set_coarse_mode(mode)
        out[mode] = tuples(det.match(img, 65))
        say("threshold 65 mode", mode, "matches", len(out[mode]), "candidates", det.last_timing()["coarse_candidates"])
    det.set_coarse_mode(1)
    assert out[0] == out[2]
    ref = oracle_tuples(oracle.match(img, [4, 8], 65, bank, nthreads=NTHREADS))
    assert out[0] == ref


# ------------------------------------------------------------------------------------------------
# 8191 features at the coarse level (the 8-orientation limit) on both coarse paths
# ------------------------------------------------------------------------------------------------
def limit_template(g, n, box, labels):
    """n features at a few hundred distinct points of a box, repeated: sums come close to 4 * n."""
    pts = np.column_stack([g.integers(0, box, 400), g.integers(0, box, 400), g.choice(labels, 400)])
    idx = np.concatenate([np.zeros(n - 400, np.int64), np.arange(400)])  # point 0 carries most of the weight
    return pts[idx].astype(np.int32)


@pytest.mark.parametrize("mode", [0, 2])
def test_8191_coarse_features_both_paths_equal_oracle(mode):
    g = np.random.default_rng(11)
    _, _, poly, _ = synth.scene_plain(0)
    img = synth.scene_sized(31, 960, 1280, poly, n_shapes=30)
    det = sbm.Detector(128, [4, 8])
    ref_bank = {}
    for k in range(6):
        f1 = limit_template(g, 8191, 45, np.arange(8))
        f0 = limit_template(g, 8191, 90, np.arange(8))
        pyr = [dict(width=90, height=90, tl_x=0, tl_y=0, features=f0),
               dict(width=45, height=45, tl_x=0, tl_y=0, features=f1)]
        det.add_template_pyramid("lim", pyr)
        ref_bank.setdefault("lim", []).append(pyr)
    # one template whose every feature sits on one point with every orientation in turn: raw reaches 4 * 8191
    # wherever that point's spread set holds the orientation
    f = np.column_stack([np.full(8191, 20), np.full(8191, 20), np.arange(8191) % 8]).astype(np.int32)
    pyr = [dict(width=40, height=40, tl_x=0, tl_y=0, features=f), dict(width=20, height=20, tl_x=0, tl_y=0, features=f)]
    det.add_template_pyramid("lim", pyr)
    ref_bank["lim"].append(pyr)
    det.set_coarse_mode(mode)
    for thr in (40, 20):
        ref = oracle_tuples(oracle.match(img, [4, 8], thr, ref_bank, nthreads=NTHREADS))
        got = tuples(det.match(img, thr))
        path = det.last_coarse_path()
        say("mode", mode, "threshold", thr, "path", path, "matches", len(got), "max score",
            max((m[2] for m in got), default=None))
        assert path[0] == (mode == 2)
        assert got == ref
        assert len(ref) > 0


# ------------------------------------------------------------------------------------------------
# 64 mixed-size scenes x 720 templates x 512 features
# ------------------------------------------------------------------------------------------------
@pytest.fixture(scope="module")
def batch64():
    imgs, poly = synth.mixed_batch(4, 64)
    bank = {"poly": [synth.bank_template(poly, a * 0.5, 1.0, 60, 512) for a in range(720)]}
    det = detector_for(bank, 512)
    dimgs = [det.upload_image(im) for im in imgs]
    alone = [tuples(det.match_image(d, 80)) for d in dimgs]
    say("64 scenes:", {s: sum(1 for im in imgs if im.shape == s) for s in {im.shape for im in imgs}},
        "matches", sum(len(a) for a in alone))
    return det, bank, imgs, dimgs, alone


def test_batch64_batched_equals_alone(batch64):
    det, _, imgs, dimgs, alone = batch64
    for slots in (4, 1):
        for rep in range(2):
            got = det.match_images(dimgs, 80, slots=slots)
            for k in range(len(imgs)):
                assert tuples(got[k]) == alone[k], f"slots {slots} pass {rep} scene {k}"
    # host inputs through the same calls
    got = det.match(imgs[0], 80)
    assert tuples(got) == alone[0]


@pytest.mark.parametrize("part", range(4))
def test_batch64_every_scene_equals_oracle(batch64, part):
    _, bank, imgs, _, alone = batch64
    for k in range(part, len(imgs), 4):
        ref = oracle_tuples(oracle.match(imgs[k], [4, 8], 80, bank, nthreads=NTHREADS))
        assert alone[k] == ref, f"scene {k} {imgs[k].shape}"
    say("part", part, "ok")


# ------------------------------------------------------------------------------------------------
# serving over the 5,000-template bank, features U[64, 2048]
# ------------------------------------------------------------------------------------------------
@pytest.fixture(scope="module")
def serve5k():
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
    det = detector_for(bank, 2048)
    assert det.numTemplates() == 5000
    alone = [tuples(det.match(im, 80)) for im in pool]
    say("serving pool:", len(pool), "scenes, matches", sum(len(a) for a in alone))
    return det, bank, pool, alone


def test_serve5k_every_request_equals_alone(serve5k):
    det, _, pool, alone = serve5k
    srv = sbm.Server(det, slots=4, hbm_budget_bytes=256 << 20)
    try:
        ids, lock = [], threading.Lock()
        rng = np.random.default_rng(9)
        order = rng.permutation(np.repeat(np.arange(len(pool)), 3))

        def client(k):
            for i in order[k::2]:
                rid = srv.submit(pool[int(i)], 80)
                with lock:
                    ids.append((rid, int(i)))

        th = [threading.Thread(target=client, args=(k,)) for k in range(2)]
        for t in th:
            t.start()
        for t in th:
            t.join()
        lat = []
        for rid, i in ids:
            got, ms = srv.wait(rid)
            assert tuples(got) == alone[i], f"request {rid} (scene {i})"
            lat.append(ms)
        st = srv.stats()
        say("served", len(ids), "requests; p50 %.2f ms p99 %.2f ms" % (np.percentile(lat, 50), np.percentile(lat, 99)),
            st)
        assert st["completed"] == len(ids) and st["failed"] == 0
        assert st["peak_bytes"] <= 256 << 20
    finally:
        srv.close()


def test_serve5k_sample_equals_oracle(serve5k):
    _, bank, pool, alone = serve5k
    for i in (0, 1, 2, 3, 4, 5):  # two scenes of each size
        ref = oracle_tuples(oracle.match(pool[i], [4, 8], 80, bank, nthreads=NTHREADS))
        assert alone[i] == ref, f"scene {i} {pool[i].shape}"
