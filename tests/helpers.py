"""Shared test helpers: fixture loading and bank conversion between the product and the oracle."""
import gzip
import json
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def golden(name):
    return os.path.join(GOLDEN, name)


def lut_table():
    return np.array(json.load(open(golden("similarity_lut.json"))), np.uint8)


def scene_bgr(case):
    from PIL import Image
    rgb = np.asarray(Image.open(golden(f"case{case}_scene.png")).convert("RGB"))
    return np.ascontiguousarray(rgb[:, :, ::-1])


def to_gray(bgr):
    """BT.601 luma in 14-bit fixed point (only used to make a gray input; not a parity target)."""
    b, g, r = (bgr[:, :, i].astype(np.int64) for i in range(3))
    return ((r * 4899 + g * 9617 + b * 1868 + 8192) >> 14).astype(np.uint8)


FIXTURE_BANKS = {0: ("case0_circle_templ.yaml.gz", "circle"), 1: ("case1_test_templ.yaml.gz", "test"),
                 2: ("case2_test_templ.yaml.gz", "test")}


def load_fixture_detector(case, num_features=63):
    import shape_based_matching_amd as sbm
    det = sbm.Detector(num_features, [4, 8])
    det.readClass(golden(FIXTURE_BANKS[case][0]))
    return det


def bank_of(det, class_ids=None):
    """{class_id: [pyramid(list of level dicts)]} read back through the product's getters."""
    out = {}
    for cid in det.classIds():
        if class_ids and cid not in class_ids:
            continue
        out[cid] = [[t.as_dict() for t in det.getTemplates(cid, i)] for i in range(det.numTemplates(cid))]
    return out


def tuples(matches):
    return [(m.x, m.y, np.float32(m.similarity), m.class_id, m.template_id) for m in matches]


def oracle_tuples(res):
    return [(x, y, np.float32(s), c, t) for (x, y, s, c, t) in res]
