"""Writes tests/golden/oracle_digests.json: SHA-256 digests of the oracle's intermediates and its full
match lists on the reference's scenes, so that the oracle and the GPU path are both held to recorded
outputs (any later change to either shows up as a digest mismatch).

Run in the build container after the oracle is built:

    python tests/golden/make_digests.py

Contents (all from the build's own oracle on the committed fixtures; no reference code runs):
  bin_table8 / bin_table16   the exhaustive (dx, dy) -> bin tables, dx, dy in [-1020, 1020]
  cases.<k>.<gray|bgr>       per pyramid level (T = {4, 8}): quantized, spread, response maps, linear memories;
                             and the match list at threshold 80 against the case's fixture bank
  bins16                     the 16-orientation variant on case 1 (BGR): a bank of 12 rotations of
                             case1_train.png (digest of its features), per-level quantized and linear
                             memories, and the match list at threshold 50
Similarities are stored as float32 bit patterns.
"""
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))

import oracle  # noqa: E402
from helpers import bank_of, load_fixture_detector, scene_bgr, to_gray  # noqa: E402

T = [4, 8]
THR = 80.0


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def crop16(img):
    r, c = img.shape[:2]
    return np.ascontiguousarray(img[: r - r % 16, : c - c % 16])


def match_rows(res):
    return [[x, y, int(np.float32(s).view(np.uint32)), c, t] for (x, y, s, c, t) in res]


def levels8(img):
    out = []
    cur = img
    lms = oracle.lm_pyramid(img, T)
    for lvl, t in enumerate(T):
        if lvl:
            cur = oracle.pyr_down(cur)
        q, _, _ = oracle.quantized_orientations(cur)
        s = oracle.spread(q, t)
        r, c = q.shape
        out.append(dict(q=sha(q), spread=sha(s), resp=sha(oracle.response_maps(s)), lm=sha(lms[lvl][: 8 * r * c])))
    return out


def train16():
    from PIL import Image
    rgb = np.asarray(Image.open(os.path.join(HERE, "case1_train.png")).convert("RGB"))
    img = np.ascontiguousarray(rgb[:, :, ::-1])
    full = np.full(img.shape[:2], 255, np.uint8)
    bank = []
    for a in range(0, 360, 30):
        pyr = oracle.add_template16(oracle.shape_transform(img, a, 1.0), 128, 2, mask=oracle.mask_of(full, a, 1.0))
        bank.append(pyr)
    return {"test": bank}


def main():
    out = {"bin_table8": sha(oracle.bin_table()), "bin_table16": sha(oracle.bin_table16()), "cases": {}}
    for case in (0, 1, 2):
        bgr = crop16(scene_bgr(case))
        bank = bank_of(load_fixture_detector(case))
        entry = {"shape": list(bgr.shape[:2])}
        for kind, img in (("gray", to_gray(bgr)), ("bgr", bgr)):
            entry[kind] = dict(levels=levels8(img), matches=match_rows(oracle.match(img, T, THR, bank, nthreads=8)))
            print("case", case, kind, len(entry[kind]["matches"]), "matches", flush=True)
        out["cases"][str(case)] = entry
    bank16 = train16()
    img = crop16(scene_bgr(1))
    feats = np.concatenate([np.asarray(lv["features"], np.int32).ravel() for p in bank16["test"] for lv in p])
    lms = oracle.lm_pyramid16(img, T)
    lv16 = []
    cur = img
    for lvl in range(len(T)):
        if lvl:
            cur = oracle.pyr_down(cur)
        q, _, _ = oracle.quantized_orientations16(cur)
        r, c = q.shape
        lv16.append(dict(q=sha(q), lm=sha(lms[lvl][: 16 * r * c])))
    m16 = match_rows(oracle.match16(img, T, 50.0, bank16, nthreads=8))
    print("16 orientations:", len(m16), "matches", flush=True)
    out["bins16"] = dict(bank=sha(feats), levels=lv16, matches=m16)
    with open(os.path.join(HERE, "oracle_digests.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
