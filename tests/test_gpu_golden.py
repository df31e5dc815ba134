"""The GPU path reproduces the recorded oracle digests of tests/golden/oracle_digests.json through the C ABI:
per-level quantized orientations and linear memories, and the full match lists, on the reference's three
scenes (gray and colour) with their fixture banks, and the 16-orientation variant trained on the GPU."""
import hashlib
import json

import numpy as np
import pytest

import shape_based_matching_amd as sbm
from helpers import golden, load_fixture_detector, scene_bgr, to_gray

pytestmark = pytest.mark.gpu
GOLD = json.load(open(golden("oracle_digests.json")))


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def crop16(img):
    r, c = img.shape[:2]
    return np.ascontiguousarray(img[: r - r % 16, : c - c % 16])


def rows(matches):
    return [[m.x, m.y, int(np.float32(m.similarity).view(np.uint32)), m.class_id, m.template_id] for m in matches]


@pytest.mark.parametrize("case", [0, 1, 2])
@pytest.mark.parametrize("kind", ["gray", "bgr"])
def test_gpu_reproduces_case_digests(case, kind):
    bgr = crop16(scene_bgr(case))
    img = to_gray(bgr) if kind == "gray" else bgr
    want = GOLD["cases"][str(case)][kind]
    det = load_fixture_detector(case)
    det.prepare_image(img)
    for lvl, w in enumerate(want["levels"]):
        assert sha(det.read_quantized(lvl)) == w["q"], lvl
        assert sha(det.read_linear_memories(lvl)) == w["lm"], lvl
    for mode in (sbm.MODE_PLAIN, sbm.MODE_BATCHED):
        assert rows(det.match(img, 80, mode=mode)) == want["matches"], mode


def test_gpu_reproduces_16_orientation_digests():
    from PIL import Image
    rgb = np.asarray(Image.open(golden("case1_train.png")).convert("RGB"))
    train = np.ascontiguousarray(rgb[:, :, ::-1])
    det = sbm.Detector16(128, [4, 8])
    assert det.add_templates_transformed(train, "test", [(a, 1.0) for a in range(0, 360, 30)]) == list(range(12))
    feats = np.concatenate([t.features.ravel() for i in range(12) for t in det.getTemplates("test", i)])
    assert sha(feats) == GOLD["bins16"]["bank"]
    img = crop16(scene_bgr(1))
    det.prepare_image(img)
    for lvl, w in enumerate(GOLD["bins16"]["levels"]):
        assert sha(det.read_quantized(lvl)) == w["q"], lvl
        assert sha(det.read_linear_memories(lvl)) == w["lm"], lvl
    assert rows(det.match(img, 50)) == GOLD["bins16"]["matches"]
