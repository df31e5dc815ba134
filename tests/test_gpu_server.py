"""Serving loop (sbm_server): every request's matches equal what the detector returns for that image
alone, whatever else is in flight, with deferrals under a tight device-memory budget and a clean
failure for a request larger than the whole budget."""
import threading

import numpy as np
import pytest

import shape_based_matching_amd as sbm
from helpers import tuples
from shape_based_matching_amd import synth

pytestmark = pytest.mark.gpu

SIZES = [(480, 640), (960, 1280), (480, 640), (1936, 2592), (960, 1280), (480, 640), (1936, 2592), (480, 640),
         (960, 1280), (480, 640)]


def say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def served():
    _, _, poly, _ = synth.scene_plain(0)
    det = sbm.Detector(128, [4, 8])
    for ang in range(0, 360, 3):
        det.add_template_pyramid("poly", synth.bank_template(poly, float(ang), 1.0, 60, 128))
    imgs = [synth.scene_sized(100 + i, r, c, poly, n_shapes=12) for i, (r, c) in enumerate(SIZES)]
    imgs[2] = np.dstack([imgs[2], imgs[2][:, ::-1], 255 - imgs[2]])  # one colour request
    masks = [None] * len(imgs)
    m = np.zeros(imgs[5].shape[:2], np.uint8)
    m[40:-40, 60:-60] = 255
    masks[5] = m  # one masked request
    alone = [tuples(det.match(im, 75, masks=mk)) for im, mk in zip(imgs, masks)]
    assert sum(len(a) for a in alone) > 0
    return det, imgs, masks, alone


def test_results_equal_alone_in_concurrent_mix(served):
    det, imgs, masks, alone = served
    srv = sbm.Server(det, slots=4, hbm_budget_bytes=8 << 30)
    try:
        say("3 submitting threads x", len(imgs), "requests, 4 slots")
        ids = {}
        lock = threading.Lock()

        def client(k):
            for rep in range(2):
                for i in range(k, len(imgs), 3):
                    rid = srv.submit(imgs[i], 75, mask=masks[i])
                    with lock:
                        ids[rid] = i

        th = [threading.Thread(target=client, args=(k,)) for k in range(3)]
        for t in th:
            t.start()
        for t in th:
            t.join()
        lat = []
        for rid, i in ids.items():
            got, ms = srv.wait(rid)
            assert tuples(got) == alone[i], f"request {rid} (image {i})"
            lat.append(ms)
        st = srv.stats()
        say("  stats", st, "latency ms p50 %.2f max %.2f" % (np.median(lat), max(lat)))
        assert st["completed"] == 2 * len(imgs) and st["failed"] == 0
        assert st["peak_bytes"] <= 8 << 30
    finally:
        srv.close()


def test_tight_budget_defers_and_keeps_results(served):
    det, imgs, masks, alone = served
    big = [i for i, (r, c) in enumerate(SIZES) if r == 1936]
    one = sbm.Server(det, slots=1, hbm_budget_bytes=8 << 30)
    try:
        got, _ = one.wait(one.submit(imgs[big[0]], 75))
        assert tuples(got) == alone[big[0]]
        need = one.stats()["peak_bytes"]
    finally:
        one.close()
    budget = int(need * 1.3)  # room for one large request and a little more, never two large ones
    srv = sbm.Server(det, slots=4, hbm_budget_bytes=budget)
    try:
        say("budget", budget, "bytes (one large request needs", need, ")")
        order = big + list(range(len(imgs))) + big
        ids = [(srv.submit(imgs[i], 75, mask=masks[i]), i) for i in order]
        for rid, i in ids:
            got, _ = srv.wait(rid)
            assert tuples(got) == alone[i], f"request {rid} (image {i})"
        st = srv.stats()
        say("  stats", st)
        assert st["completed"] == len(order) and st["failed"] == 0
        assert st["deferrals"] > 0
        assert st["peak_bytes"] <= budget
    finally:
        srv.close()


def test_request_over_budget_fails_alone(served):
    det, imgs, masks, alone = served
    small = [i for i, (r, c) in enumerate(SIZES) if r == 480 and imgs[i].ndim == 2 and masks[i] is None]
    big = [i for i, (r, c) in enumerate(SIZES) if r == 1936][0]
    one = sbm.Server(det, slots=1, hbm_budget_bytes=8 << 30)
    try:
        one.wait(one.submit(imgs[small[0]], 75))
        need_small = one.stats()["peak_bytes"]
    finally:
        one.close()
    srv = sbm.Server(det, slots=2, hbm_budget_bytes=int(need_small * 1.5))
    try:
        a = srv.submit(imgs[small[0]], 75)
        b = srv.submit(imgs[big], 75)
        c = srv.submit(imgs[small[1]], 75)
        assert tuples(srv.wait(a)[0]) == alone[small[0]]
        with pytest.raises(sbm.SbmError) as e:
            srv.wait(b)
        assert e.value.code == -6
        assert tuples(srv.wait(c)[0]) == alone[small[1]]
        # malformed requests are refused at submission, queued work is unaffected
        with pytest.raises(sbm.SbmError) as e:
            srv.submit(np.zeros((1000, 1000), np.uint8), 75)
        assert e.value.code == -2
        with pytest.raises(sbm.SbmError) as e:
            srv.wait(10 ** 9, timeout_ms=5)
        assert e.value.code == -11
        st = srv.stats()
        assert st["completed"] == 2 and st["failed"] == 1
    finally:
        srv.close()


def test_server_started_from_a_packed_bank(served, tmp_path):
    """A serving detector loaded from the packed bank file (no YAML parsing) returns the same matches."""
    det, imgs, masks, alone = served
    det.save_bank(tmp_path / "served.bank")
    fresh = sbm.Detector(128, [4, 8])
    assert fresh.load_bank(tmp_path / "served.bank") == 1
    fresh.upload_bank()  # the start-up pays the bank's upload, not the first request
    srv = sbm.Server(fresh, slots=2)
    try:
        ids = [srv.submit(im, 75, mask=mk) for im, mk in zip(imgs[:4], masks[:4])]
        for k, i in enumerate(ids):
            assert tuples(srv.wait(i)[0]) == alone[k], k
    finally:
        srv.close()


def test_queued_requests_share_launches(served):
    """Requests of one size, threshold and selection queued together go out in shared launches (every kernel
    covering all of them); each result still equals the image matched alone, masks and colour included."""
    det, imgs, masks, alone = served
    srv = sbm.Server(det, slots=2, hbm_budget_bytes=8 << 30)
    try:
        order = [i for i, (r, c) in enumerate(SIZES) if r == 480] * 4 + [i for i, (r, c) in enumerate(SIZES)
                                                                          if r == 960] * 3
        ids = [(srv.submit(imgs[i], 75, mask=masks[i]), i) for i in order]
        for rid, i in ids:
            got, _ = srv.wait(rid)
            assert tuples(got) == alone[i], f"request {rid} (image {i})"
        st = srv.stats()
        say("  stats", st)
        assert st["completed"] == len(order) and st["failed"] == 0
        assert st["launches"] < st["admitted"]  # some launches carried several requests
    finally:
        srv.close()


def test_budget_counts_the_bank(served):
    """The budget covers the resident bank as well as the slots: peak_bytes is at least the bank's bytes, and a
    budget below the bank alone fails every request with SBM_EBUDGET."""
    det, imgs, masks, alone = served
    small = [i for i, (r, c) in enumerate(SIZES) if r == 480 and imgs[i].ndim == 2 and masks[i] is None][0]
    feats = sum(int(lv.features.shape[0]) for t in range(det.numTemplates("poly")) for lv in det.getTemplates("poly", t))
    bank_bytes = 8 * feats  # packed features alone, the records and indices come on top
    srv = sbm.Server(det, slots=1, hbm_budget_bytes=8 << 30)
    try:
        assert tuples(srv.wait(srv.submit(imgs[small], 75))[0]) == alone[small]
        assert srv.stats()["peak_bytes"] > bank_bytes
    finally:
        srv.close()
    srv = sbm.Server(det, slots=1, hbm_budget_bytes=bank_bytes // 2)
    try:
        with pytest.raises(sbm.SbmError) as e:
            srv.wait(srv.submit(imgs[small], 75))
        assert e.value.code == -6
    finally:
        srv.close()
    assert tuples(det.match(imgs[small], 75)) == alone[small]  # the detector's own buffers grow again
