"""match followed by the demos' NMS step, on the device (sbm_match_nms / sbm_match_image_nms).

The kept list must equal cv_dnn::NMSBoxes (the oracle's restatement, test.cpp:31-121, and the
library's host nms_boxes) applied to Detector::match's list with the demos' boxes
(x, y, templ[0].width, templ[0].height) and eta 1 (test.cpp:428-451); the full list returned beside
it must equal match's.  Lists longer than the device kernel holds take the host path and must give
the same answer.
"""
import os

import numpy as np
import pytest

import oracle
import shape_based_matching_amd as sbm
from helpers import bank_of, oracle_tuples, tuples
from shape_based_matching_amd import synth

pytestmark = pytest.mark.gpu
NTHREADS = max(1, min(16, os.cpu_count() or 1))
DEVICE_LIST_MAX = 2048


def say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def scene():
    img, _, poly, _ = synth.scene_plain(0)
    det = sbm.Detector(128, [4, 8])
    for ang in range(0, 360, 2):
        det.add_template_pyramid("poly", synth.bank_template(poly, float(ang), 1.0, 60, 128))
    for ang in range(0, 360, 6):
        det.add_template_pyramid("small", synth.bank_template(poly, float(ang), 0.7, 60, 96))
    return det, img


def reference_kept(det, matches, score_threshold, nms_threshold, top_k):
    sizes = {}
    boxes, scores = [], []
    for m in matches:
        key = (m.class_id, m.template_id)
        if key not in sizes:
            t0 = det.getTemplates(m.class_id, m.template_id)[0]
            sizes[key] = (t0.width, t0.height)
        boxes.append((m.x, m.y) + sizes[key])
        scores.append(m.similarity)
    idx = oracle.nms_boxes(boxes, scores, score_threshold, nms_threshold, 1.0, top_k)
    assert idx == sbm.nms_boxes(boxes, scores, score_threshold, nms_threshold, 1.0, top_k)
    return [matches[i] for i in idx]


def rows(xyct_sim_names):
    xyct, sim, names = xyct_sim_names
    return [(int(r[0]), int(r[1]), np.float32(s), names[r[2]], int(r[3])) for r, s in zip(xyct, sim)]


@pytest.mark.parametrize("threshold,score_threshold,nms_threshold,top_k", [
    (90, 0.0, 0.5, 0), (75, 0.0, 0.5, 0), (75, 85.0, 0.3, 0), (75, 75.0, 0.0, 0), (75, 0.0, 0.5, 7),
    (60, 0.0, 0.5, 0), (60, 70.0, 0.8, 3), (99.9, 0.0, 0.5, 0)])
def test_kept_equals_nms_of_match(scene, threshold, score_threshold, nms_threshold, top_k):
    det, img = scene
    say("640x480, 240 templates, threshold", threshold, "nms", score_threshold, nms_threshold, top_k)
    ref_all = det.match(img, threshold)
    ref = tuples(reference_kept(det, ref_all, score_threshold, nms_threshold, top_k))
    kept, full = det.match_nms_arrays(img, threshold, score_threshold=score_threshold, nms_threshold=nms_threshold,
                                      top_k=top_k, with_all=True)
    assert rows(full) == tuples(ref_all)
    assert rows(kept) == ref
    assert tuples(det.match_nms(img, threshold, score_threshold=score_threshold, nms_threshold=nms_threshold,
                                top_k=top_k)) == ref
    say("  list", len(ref_all), "kept", len(ref))
    if threshold == 75:
        assert 0 < len(ref) < len(ref_all) < DEVICE_LIST_MAX  # the device kernel's path, with suppression


def test_full_list_equals_oracle(scene):
    det, img = scene
    ref = oracle_tuples(oracle.match(img, [4, 8], 80, bank_of(det), nthreads=NTHREADS))
    kept, full = det.match_nms_arrays(img, 80, with_all=True)
    assert rows(full) == ref
    assert len(ref) > 0


def test_resident_colour_masked_and_selected(scene):
    det, img = scene
    bgr = np.ascontiguousarray(np.dstack([img, img[:, ::-1], 255 - img]))
    mask = np.zeros(img.shape, np.uint8)
    mask[30:-50, 40:-20] = 255
    for src, mk in ((img, mask), (bgr, None), (bgr, mask)):
        for ids in (None, ["small"], ["poly", "small"], []):
            say("shape", src.shape, "mask", mk is not None, "classes", ids)
            ref_all = det.match(src, 70, class_ids=ids, masks=mk)
            ref = tuples(reference_kept(det, ref_all, 0.0, 0.4, 0))
            assert tuples(det.match_nms(src, 70, class_ids=ids, nms_threshold=0.4, mask=mk)) == ref
            dimg = det.upload_image(src, mk)
            assert tuples(det.match_nms(dimg, 70, class_ids=ids, nms_threshold=0.4)) == ref


def test_long_list_takes_host_path_with_same_result(scene):
    det, img = scene
    for threshold in (50, 40, 30, 20):
        ref_all = det.match(img, threshold)
        if len(ref_all) > DEVICE_LIST_MAX:
            break
    say("threshold", threshold, "list", len(ref_all))
    assert len(ref_all) > DEVICE_LIST_MAX
    ref = tuples(reference_kept(det, ref_all, 0.0, 0.5, 0))
    kept, full = det.match_nms_arrays(img, threshold, nms_threshold=0.5, with_all=True)
    assert rows(full) == tuples(ref_all)
    assert rows(kept) == ref
    # and back to a short list on the same detector
    ref_all = det.match(img, 80)
    assert tuples(det.match_nms(img, 80)) == tuples(reference_kept(det, ref_all, 0.0, 0.5, 0))
