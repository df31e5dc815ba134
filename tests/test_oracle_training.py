"""Pins the oracle's training path (extraction, warpAffine, produce_infos) against the reference's own
trained banks.

The three fixture banks were written by the reference's demos (test.cpp):

* case1 (angle_test, test.cpp:224-292): addTemplate on a padded crop of train.png at angle 0, then
  addTemplate_rotate for 1..360 degrees.  It was made by the current line2Dup.cpp, so extraction and
  rotation must reproduce it exactly, labels included (the labels of the rotated templates depend on
  each feature's float orientation, i.e. on fastAtan2).
* case2 (noise_test, test.cpp:381-410) and case0 (scale_test, test.cpp:126-162): addTemplate on
  shapeInfo_producer images (warpAffine) at 361 angles / 89 scales.  These files predate the current
  extractTemplate: they follow the original linemod feature selection (stop at exactly num_features),
  refuse a level with fewer candidates than num_features, and rotate about (cols/2, rows/2) in
  integer arithmetic.  The tests restate those three older rules here and keep everything else from
  the oracle, so they pin the oracle's warpAffine and candidate stage (gradients, 5x5 NMS, strong
  threshold, stable sort) against the reference's real output.  All level-0 templates match; a few
  level-1 templates (listed below) differ in one or two near-tied features.
"""
import gzip
import re

import numpy as np
import pytest

import oracle
from helpers import golden, load_fixture_detector

# level-1 templates of the older fixture banks that differ in near-tied features (see module doc)
CASE2_LEVEL1_KNOWN = {103, 124, 162, 271, 304, 342}
CASE0_LEVEL1_KNOWN = {0.44, 0.66}


def png_bgr(name):
    from PIL import Image
    rgb = np.asarray(Image.open(golden(name)).convert("RGB"))
    return np.ascontiguousarray(rgb[:, :, ::-1])


def fixture_infos(name):
    text = gzip.open(golden(name)).read().decode()
    a = [np.float32(float(v)) for v in re.findall(r"angle: ([-0-9.e+]+)", text)]
    s = [np.float32(float(v)) for v in re.findall(r"scale: ([-0-9.e+]+)", text)]
    return list(zip(a, s))


def select_linemod_original(cands_xy, num_features, distance):
    """The selection the older fixture banks were made with: keep scanning the score-sorted
    candidates, relaxing the distance by 1 per pass, until exactly num_features are kept."""
    feats, dsq, i = [], distance * distance, 0
    while len(feats) < num_features:
        x, y = cands_xy[i]
        if all((x - f[0]) ** 2 + (y - f[1]) ** 2 >= dsq for f in feats):
            feats.append((x, y, i))
        i += 1
        if i == len(cands_xy):
            i, distance = 0, distance - 1.0
            dsq = distance * distance
    return [f[2] for f in feats]


def train_older(img, mask, num_features, levels=2):
    """addTemplate as the older fixture banks were made (selection and refusal rule above)."""
    feats = []
    cur, cm, nf = img, mask, num_features
    for lvl in range(levels):
        if lvl:
            nf //= 2
            cur, cm = oracle.pyr_down(cur), oracle.mask_down(cm)
        q, mag, ang = oracle.quantized_orientations(cur)
        xyl, _, _ = oracle.extract_candidates(mag, q, ang, cm)
        if len(xyl) < nf:
            return None
        idx = select_linemod_original([tuple(p) for p in xyl[:, :2].tolist()], nf, float(len(xyl) // nf + 1))
        feats.append(xyl[idx])
    return crop(feats)


def crop(levels):
    """cropTemplates (line2Dup.cpp:103-149) on absolute per-level features."""
    mnx = min(int((f[:, 0] << l).min()) for l, f in enumerate(levels))
    mny = min(int((f[:, 1] << l).min()) for l, f in enumerate(levels))
    mxx = max(int((f[:, 0] << l).max()) for l, f in enumerate(levels))
    mxy = max(int((f[:, 1] << l).max()) for l, f in enumerate(levels))
    mnx -= mnx % 2
    mny -= mny % 2
    out = []
    for lvl, f in enumerate(levels):
        g = f.copy()
        g[:, 0] -= mnx >> lvl
        g[:, 1] -= mny >> lvl
        out.append(dict(width=(mxx - mnx) >> lvl, height=(mxy - mny) >> lvl, tl_x=mnx >> lvl, tl_y=mny >> lvl,
                        features=g))
    return out


def same_level(a, t):
    return (np.array_equal(a["features"], t.features)
            and (a["width"], a["height"], a["tl_x"], a["tl_y"]) == (t.width, t.height, t.tl_x, t.tl_y))


# ---------------------------------------------------------------------------------------------
# case1: current extraction + rotation, exact
# ---------------------------------------------------------------------------------------------
def case1_source():
    bgr = png_bgr("case1_train.png")[110:110 + 270, 130:130 + 270]  # Rect(130, 110, 270, 270)
    pad = np.zeros((470, 470, 3), np.uint8)
    pad[100:370, 100:370] = bgr
    mask = np.zeros((470, 470), np.uint8)
    mask[100:370, 100:370] = 255
    return pad, mask


def test_case1_extraction_reproduces_fixture_template0():
    pad, mask = case1_source()
    # angle 0, scale 1: shapeInfo_producer's warp is the identity
    assert np.array_equal(oracle.shape_transform(pad, 0.0, 1.0), pad)
    pyr = oracle.add_template(pad, 128, 2, mask=oracle.mask_of(mask, 0.0, 1.0))
    det = load_fixture_detector(1)
    fx = det.getTemplates("test", 0)
    for lvl in range(2):
        assert same_level(pyr[lvl], fx[lvl]), lvl


def test_case1_rotated_labels_pin_the_orientation_angles():
    """Rotating the extracted template (with the oracle's float orientations) by 1..360 degrees
    about (235, 235) reproduces all 720 template levels, labels included."""
    pad, mask = case1_source()
    pyr = oracle.add_template(pad, 128, 2, mask=mask)
    det = load_fixture_detector(1)
    for k in range(1, 361):
        rot = oracle.rotate_template(pyr, float(k), (235.0, 235.0))
        fx = det.getTemplates("test", k)
        for lvl in range(2):
            assert np.array_equal(rot[lvl]["features"], fx[lvl].features), (k, lvl)


def test_produce_infos_reproduces_fixture_angle_lists():
    for name in ("case1_test_info.yaml.gz", "case2_test_info.yaml.gz"):
        got = [(np.float32(a), np.float32(s)) for a, s in oracle.produce_infos((0, 360), (1.0,), 1.0)]
        assert got == fixture_infos(name), name
    # the scale sweep of case0 accumulates in float; the fixture lists the scales whose template
    # was accepted, in order and bit for bit
    got = [(np.float32(a), np.float32(s)) for a, s in oracle.produce_infos((), (0.1, 1.0), 15.0, 0.01)]
    fi = fixture_infos("case0_circle_info.yaml.gz")
    assert len(got) == 91 and len(fi) == 89
    it = iter(got)
    assert all(any(g == f for g in it) for f in fi)
    assert got[-1] == fi[-1] and got[-1][1] < 1  # 0.1 + 90 float steps of 0.01 falls just short of 1


# ---------------------------------------------------------------------------------------------
# warpAffine closed forms
# ---------------------------------------------------------------------------------------------
def test_warp_identity_and_quarter_turns_are_exact_permutations():
    rng = np.random.default_rng(3)
    n = 48
    img = rng.integers(0, 256, (n, n, 3), dtype=np.uint8)
    assert np.array_equal(oracle.shape_transform(img, 0.0, 1.0), img)
    # 90 deg about (n/2, n/2): dst(x, y) = src(n - y, x); the column x = n lies outside -> 0
    r90 = oracle.shape_transform(img, 90.0, 1.0)
    exp = np.zeros_like(img)
    for y in range(n):
        for x in range(n):
            sx, sy = n - y, x
            if 0 <= sx < n and 0 <= sy < n:
                exp[y, x] = img[sy, sx]
    assert np.array_equal(r90, exp)
    # 180 deg: dst(x, y) = src(n - x, n - y)
    r180 = oracle.shape_transform(img, 180.0, 1.0)
    exp = np.zeros_like(img)
    exp[1:, 1:] = img[::-1, ::-1][:-1, :-1]
    assert np.array_equal(r180, exp)


def test_warp_half_pixel_scale_is_round_half_up_average():
    """Scale 2 about (0, 0) samples src at (x/2, y/2): odd positions are the average of two or four
    pixels, rounded half up -- the 15-bit weights (32-f)*(32-g)*32 sum to 2^15 exactly."""
    rng = np.random.default_rng(4)
    img = rng.integers(0, 256, (20, 24), dtype=np.uint8)
    M = np.array([[2.0, 0.0, 0.0], [0.0, 2.0, 0.0]])
    out = oracle.warp_affine(img, M)
    a = img.astype(np.int64)
    pad = np.zeros((21, 25), np.int64)
    pad[:20, :24] = a
    for y in range(20):
        for x in range(24):
            y0, x0 = y // 2, x // 2
            if y % 2 == 0 and x % 2 == 0:
                e = pad[y0, x0]
            elif y % 2 == 0:
                e = (pad[y0, x0] + pad[y0, x0 + 1] + 1) // 2
            elif x % 2 == 0:
                e = (pad[y0, x0] + pad[y0 + 1, x0] + 1) // 2
            else:
                e = (pad[y0, x0] + pad[y0, x0 + 1] + pad[y0 + 1, x0] + pad[y0 + 1, x0 + 1] + 2) // 4
            assert out[y, x] == e, (y, x)


def test_rotation_matrix_and_inverse():
    M = oracle.rotation_matrix2d((73.5, 74.0), 30.0, 0.5)
    c, s = np.cos(np.radians(30.0)) * 0.5, np.sin(np.radians(30.0)) * 0.5
    assert np.allclose(M, [[c, s, (1 - c) * 73.5 - s * 74.0], [-s, c, s * 73.5 + (1 - c) * 74.0]], rtol=0, atol=1e-12)
    iM = oracle.invert_affine(M)
    full = np.vstack([M, [0, 0, 1]]) @ np.vstack([iM, [0, 0, 1]])
    assert np.allclose(full, np.eye(3), atol=1e-12)


# ---------------------------------------------------------------------------------------------
# case2 / case0: warpAffine + candidate stage against the older fixture banks
# ---------------------------------------------------------------------------------------------
def test_case2_warped_training_reproduces_fixture():
    bgr = png_bgr("case2_train.png")  # 148 x 147
    rows, cols = bgr.shape[:2]
    mask = np.full((rows, cols), 255, np.uint8)
    det = load_fixture_detector(2)
    infos = fixture_infos("case2_test_info.yaml.gz")
    assert det.numTemplates("test") == len(infos) == 361
    level1_diff = set()
    for k, (a, s) in enumerate(infos):
        M = oracle.rotation_matrix2d((cols // 2, rows // 2), float(a), float(s))
        m = np.where(oracle.warp_affine(mask, M) > 0, 255, 0).astype(np.uint8)
        pyr = train_older(oracle.warp_affine(bgr, M), m, 30)
        fx = det.getTemplates("test", k)
        assert same_level(pyr[0], fx[0]), k
        if not same_level(pyr[1], fx[1]):
            level1_diff.add(k)
    assert level1_diff == CASE2_LEVEL1_KNOWN


def test_case0_scaled_training_reproduces_fixture():
    bgr = png_bgr("case0_circle.png")  # 800 x 800
    mask = np.full(bgr.shape[:2], 255, np.uint8)
    det = load_fixture_detector(0)
    fi = fixture_infos("case0_circle_info.yaml.gz")
    accepted, level1_diff, k = [], set(), 0
    for a, s in oracle.produce_infos((), (0.1, 1.0), 15.0, 0.01):
        nf = int(np.float32(150) * np.float32(s))  # int(num_feature * info.scale), test.cpp:146
        pyr = train_older(oracle.shape_transform(bgr, a, s), oracle.mask_of(mask, a, s), nf)
        if pyr is None:
            continue
        accepted.append((np.float32(a), np.float32(s)))
        fx = det.getTemplates("circle", k)
        assert same_level(pyr[0], fx[0]), s
        if not same_level(pyr[1], fx[1]):
            level1_diff.add(round(float(s), 2))
        k += 1
    assert accepted == fi
    assert level1_diff == CASE0_LEVEL1_KNOWN
