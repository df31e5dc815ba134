"""The oracle still produces the recorded digests of tests/golden/oracle_digests.json (bin tables, every
intermediate per level, full match lists on the reference's scenes, and the 16-orientation variant):
a change to the checker shows up here even where the GPU path changed the same way."""
import json
import os
import sys

import pytest

from helpers import golden

sys.path.insert(0, os.path.dirname(golden("make_digests.py")))
import make_digests as md  # noqa: E402

import oracle  # noqa: E402
from helpers import bank_of, load_fixture_detector, scene_bgr, to_gray  # noqa: E402

GOLD = json.load(open(golden("oracle_digests.json")))


def test_bin_tables():
    assert md.sha(oracle.bin_table()) == GOLD["bin_table8"]
    assert md.sha(oracle.bin_table16()) == GOLD["bin_table16"]


@pytest.mark.parametrize("case", [0, 1, 2])
@pytest.mark.parametrize("kind", ["gray", "bgr"])
def test_case_intermediates_and_matches(case, kind):
    bgr = md.crop16(scene_bgr(case))
    img = to_gray(bgr) if kind == "gray" else bgr
    want = GOLD["cases"][str(case)][kind]
    assert md.levels8(img) == want["levels"]
    res = oracle.match(img, md.T, md.THR, bank_of(load_fixture_detector(case)), nthreads=8)
    assert md.match_rows(res) == want["matches"]


def test_16_orientation_variant():
    import numpy as np
    bank = md.train16()
    feats = np.concatenate([np.asarray(lv["features"], np.int32).ravel() for p in bank["test"] for lv in p])
    assert md.sha(feats) == GOLD["bins16"]["bank"]
    img = md.crop16(scene_bgr(1))
    assert md.match_rows(oracle.match16(img, md.T, 50.0, bank, nthreads=8)) == GOLD["bins16"]["matches"]
