"""Writes the golden fixtures under tests/golden/ from the reference's own data files.

Run in the build container only (it reads /root/reference, which the GPU host does not have):

    python tests/golden/make_fixtures.py

Outputs (all data, no reference source text):
  similarity_lut.json        the 256 values of SIMILARITY_LUT (line2Dup.cpp:608), parsed from the
                             source as numbers
  similarity_lut16.json      the 16-orientation variant's tables, 1024 values each, parsed as numbers:
                             "compiled" from test/ori_16bit_experiment/line2Dup_16bit_ori.cpp:575-609
                             (the one the variant uses), "generated" from LUT16.txt (LUT_gen.cpp's output)
  case0_circle_templ.yaml.gz the reference's trained banks (test/case0/circle_templ.yaml,
  case1_test_templ.yaml.gz   test/case1/test_templ.yaml, test/case2/test_templ.yaml), gzip-compressed
  case2_test_templ.yaml.gz
  case1_test_info.yaml.gz    shapeInfo_producer infos of case1 (angle/scale per template)
  case2_test_info.yaml.gz    shapeInfo_producer infos of case2 and case0 (test/case2/test_info.yaml,
  case0_circle_info.yaml.gz  test/case0/circle_info.yaml)
  case1_train.png            the training images the reference's trained banks came from, decoded
  case2_train.png            (test/case1/train.png, test/case2/train.png, test/case0/templ/circle.png),
  case0_circle.png           RGB PNG; alpha dropped as imread's default colour mode drops it
  case0_scene.png            decoded scenes, prepared as the demos do (test.cpp:174-181, 305-315,
  case1_scene.png            420-425): case0/1.jpg cropped to 896x576, case1/test.png padded by 250 and
  case2_scene.png            cropped to 1088x960, case2/test.png cropped to 1440x528; RGB PNG (the
                             tests flip to BGR, the channel order imread produces)
"""
import gzip
import json
import os
import re
import shutil

import numpy as np
from PIL import Image

REF = "/root/reference"
OUT = os.path.dirname(os.path.abspath(__file__))


def lut():
    text = open(os.path.join(REF, "line2Dup.cpp")).read()
    m = re.search(r"SIMILARITY_LUT\[256\]\s*=\s*\{([^}]*)\}", text)
    vals = [3 if v.strip() == "LUT3" else int(v) for v in m.group(1).split(",")]
    assert len(vals) == 256
    json.dump(vals, open(os.path.join(OUT, "similarity_lut.json"), "w"))


def lut16():
    def table(path):
        text = open(os.path.join(REF, path)).read()
        m = re.search(r"SIMILARITY_LUT\[1024\]\s*=\s*\{([^}]*)\}", text)
        vals = [int(v) for v in m.group(1).split(",") if v.strip()]
        assert len(vals) == 1024
        return vals

    json.dump({"compiled": table("test/ori_16bit_experiment/line2Dup_16bit_ori.cpp"),
               "generated": table("test/ori_16bit_experiment/LUT16.txt")},
              open(os.path.join(OUT, "similarity_lut16.json"), "w"))


def gz(src, dst):
    with open(os.path.join(REF, src), "rb") as f, open(os.path.join(OUT, dst), "wb") as raw, \
            gzip.GzipFile(dst[:-3], "wb", 9, raw, mtime=0) as g:  # deterministic bytes
        shutil.copyfileobj(f, g)


def rgb(path):
    return np.asarray(Image.open(os.path.join(REF, path)).convert("RGB"))


def crop_to(a, stride):
    return a[: (a.shape[0] // stride) * stride, : (a.shape[1] // stride) * stride]


def scenes():
    c0 = crop_to(rgb("test/case0/1.jpg"), 32)
    t1 = rgb("test/case1/test.png")
    pad = np.zeros((t1.shape[0] + 500, t1.shape[1] + 500, 3), np.uint8)
    pad[250:250 + t1.shape[0], 250:250 + t1.shape[1]] = t1
    c1 = crop_to(pad, 16)
    c2 = crop_to(rgb("test/case2/test.png"), 16)
    for name, a in (("case0_scene.png", c0), ("case1_scene.png", c1), ("case2_scene.png", c2)):
        Image.fromarray(np.ascontiguousarray(a)).save(os.path.join(OUT, name), optimize=True)
        print(name, a.shape)


def trains():
    for src, name in (("test/case1/train.png", "case1_train.png"), ("test/case2/train.png", "case2_train.png"),
                      ("test/case0/templ/circle.png", "case0_circle.png")):
        a = rgb(src)
        Image.fromarray(np.ascontiguousarray(a)).save(os.path.join(OUT, name), optimize=True)
        print(name, a.shape)


if __name__ == "__main__":
    lut()
    lut16()
    gz("test/case0/circle_templ.yaml", "case0_circle_templ.yaml.gz")
    gz("test/case1/test_templ.yaml", "case1_test_templ.yaml.gz")
    gz("test/case2/test_templ.yaml", "case2_test_templ.yaml.gz")
    gz("test/case1/test_info.yaml", "case1_test_info.yaml.gz")
    gz("test/case2/test_info.yaml", "case2_test_info.yaml.gz")
    gz("test/case0/circle_info.yaml", "case0_circle_info.yaml.gz")
    scenes()
    trains()
