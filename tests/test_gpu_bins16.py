"""The 16-orientation variant (test/ori_16bit_experiment/line2Dup_16bit_ori.cpp) on the GPU.

Every stage is compared with the 16-orientation oracle on the same inputs, bit for bit: quantization
(ramps at all 32 bin centres and colour scenes), masked levels and their 16-plane linear memories,
training (features of every level), and match lists on the reference's scenes and on synthetic banks
through every coarse path (gather kernels, nibble-packed gather, the nonzero-response list), plain mode,
captured graphs and batches.  The variant's own rules are checked too: 3-channel images only, no
addTemplate_rotate, up to 16383 features per level.
"""
import os

import numpy as np
import pytest

import oracle
import shape_based_matching_amd as sbm
from helpers import bank_of, golden, oracle_tuples, scene_bgr, tuples
from shape_based_matching_amd import synth

pytestmark = pytest.mark.gpu
NTHREADS = max(1, min(16, os.cpu_count() or 1))


def say(*a):
    print(*a, flush=True)


def bgr_of(gray, seed):
    """A colour image whose channels differ, so that the strongest-channel choice matters."""
    g = np.random.default_rng(seed)
    b = gray.astype(np.int32)
    gr = np.clip(np.roll(b, 2, axis=1) * 0.8 + g.normal(0, 4, b.shape), 0, 255)
    r = np.clip(255 - np.roll(b, 3, axis=0), 0, 255)
    return np.ascontiguousarray(np.stack([b, gr, r], -1).astype(np.uint8))


def train_image(name):
    from PIL import Image
    rgb = np.asarray(Image.open(golden(name)).convert("RGB"))
    return np.ascontiguousarray(rgb[:, :, ::-1])


def crop(img, m):
    r, c = img.shape[:2]
    return np.ascontiguousarray(img[: r - r % m, : c - c % m])


def ramp(theta_deg, n=48, slope=3.0):
    yy, xx = np.mgrid[0:n, 0:n].astype(np.float64)
    th = np.deg2rad(theta_deg)
    g = np.clip(np.round(128 + slope * (np.cos(th) * (xx - n / 2) + np.sin(th) * (yy - n / 2))), 0, 255)
    return np.ascontiguousarray(np.stack([g, g, g], -1).astype(np.uint8))


def bank16(poly, n_templ, n0, seed=0):
    """Synthetic two-level pyramids with labels spread over the 16 orientations."""
    g = np.random.default_rng(seed)
    out = []
    for k in range(n_templ):
        pyr = synth.bank_template(poly, k * 360.0 / n_templ, 0.9 + 0.1 * (k % 3), 60, n0)
        for lv in pyr:
            f = lv["features"].copy()
            f[:, 2] = 2 * f[:, 2] + g.integers(0, 2, len(f))
            lv["features"] = f
        out.append(pyr)
    return out


# ---------------------------------------------------------------------------------------------
# quantization and levels
# ---------------------------------------------------------------------------------------------
def test_gradient16_equals_oracle():
    det = sbm.Detector16([4, 8])
    say("ramps")
    for k in range(32):
        img = ramp(k * 11.25)
        q, mag, _ = det.gradient(img)
        q0, mag0, _ = oracle.quantized_orientations16(img)
        assert q.dtype == np.uint16 and np.array_equal(q, q0), k
        assert np.array_equal(mag, mag0)
        assert (q[10:-10, 10:-10] == (1 << (k % 16))).all()
    for seed, shape in [(0, (96, 128)), (1, (131, 197)), (2, (256, 320))]:
        say("scene", seed, shape)
        gray = synth.scene_plain(seed, rows=shape[0], cols=shape[1], n_shapes=6)[0]
        img = bgr_of(gray, seed)
        q, mag, _ = det.gradient(img)
        q0, mag0, _ = oracle.quantized_orientations16(img)
        assert np.array_equal(q, q0) and np.array_equal(mag, mag0)
        assert len(np.unique(q)) == 17  # all sixteen orientations and none


@pytest.mark.parametrize("T", [[4, 8], [5, 8], [2, 4, 8], [16]])
def test_levels16_equal_oracle(T):
    img = np.ascontiguousarray(scene_bgr(1)[:480, :640])  # divisible at every level of each pyramid here
    rows, cols = img.shape[:2]
    yy, xx = np.mgrid[0:rows, 0:cols]
    mask = (((xx - cols / 2) ** 2 + (yy - rows / 2) ** 2) < (0.4 * rows) ** 2).astype(np.uint8) * 255
    det = sbm.Detector16(T)
    say("T", T, img.shape)
    det.prepare_image(img, mask)
    lms = oracle.lm_pyramid16(img, T, mask=mask)
    r, c, mk = rows, cols, mask
    cur = img
    for lvl, t in enumerate(T):
        if lvl > 0:
            r, c = r // 2, c // 2
            cur = oracle.pyr_down(cur)
            mk = oracle.mask_down(mk)
        q0, _, _ = oracle.quantized_orientations16(cur)
        q0 = np.where(mk > 0, q0, 0).astype(np.uint16)
        assert np.array_equal(det.read_quantized(lvl), q0), lvl
        got = det.read_linear_memories(lvl)
        assert np.array_equal(got, lms[lvl][: 16 * r * c]), lvl
        assert set(np.unique(got)) <= {0, 1, 4}


# ---------------------------------------------------------------------------------------------
# training
# ---------------------------------------------------------------------------------------------
@pytest.mark.parametrize("name,nf", [("case1_train.png", 128), ("case2_train.png", 40), ("case0_circle.png", 63)])
def test_training16_equals_oracle(name, nf):
    img = train_image(name)
    det = sbm.Detector16(nf, [4, 8])
    tid = det.addTemplate(img, "a", None, nf)
    ref = oracle.add_template16(img, nf, 2)
    assert tid == 0 and ref is not None
    got = det.getTemplates("a", 0)
    for lvl in range(2):
        assert np.array_equal(got[lvl].features, ref[lvl]["features"]), lvl
        assert (got[lvl].width, got[lvl].height, got[lvl].tl_x, got[lvl].tl_y) == \
               (ref[lvl]["width"], ref[lvl]["height"], ref[lvl]["tl_x"], ref[lvl]["tl_y"])
    assert len(got[0].features) == nf and len(got[1].features) == nf // 2
    # no exhaustive mode: more features than candidates refuses the template
    assert det.addTemplate(img, "a", None, 5000) == -1 and oracle.add_template16(img, 5000, 2) is None


def test_transformed_training16_equals_oracle():
    img = train_image("case1_train.png")
    rows, cols = img.shape[:2]
    det = sbm.Detector16(100, [4, 8])
    infos = [(a, 1.0) for a in range(0, 360, 45)] + [(10.0, 0.8)]
    ids = det.add_templates_transformed(img, "t", infos)
    for k, (a, s) in enumerate(infos):
        ref = oracle.add_template16(oracle.shape_transform(img, a, s), 100, 2,
                                    mask=oracle.mask_of(np.full((rows, cols), 255, np.uint8), a, s))
        if ref is None:
            assert ids[k] == -1
            continue
        got = det.getTemplates("t", ids[k])
        for lvl in range(2):
            assert np.array_equal(got[lvl].features, ref[lvl]["features"]), (k, lvl)


# ---------------------------------------------------------------------------------------------
# matching
# ---------------------------------------------------------------------------------------------
def case_detector(nf=128):
    img = train_image("case1_train.png")
    det = sbm.Detector16(nf, [4, 8])
    ids = det.add_templates_transformed(img, "test", [(a, 1.0) for a in range(0, 360, 30)])
    assert all(i >= 0 for i in ids)
    return det


@pytest.mark.parametrize("case,thr", [(0, 35), (1, 50), (2, 35)])
def test_match16_reference_scenes_equal_oracle(case, thr):
    det = case_detector()
    img = crop(scene_bgr(case), 16)
    ref = oracle_tuples(oracle.match16(img, [4, 8], thr, bank_of(det), nthreads=NTHREADS))
    for mode in (sbm.MODE_PLAIN, sbm.MODE_BATCHED):
        got = tuples(det.match(img, thr, mode=mode))
        say("case", case, "mode", mode, "matches", len(got), "oracle", len(ref))
        assert got == ref
    assert len(ref) > 10


def test_match16_every_coarse_path_equals_oracle():
    _, _, poly, _ = synth.scene_plain(0)
    det = sbm.Detector16(128, [4, 8])
    for k, pyr in enumerate(bank16(poly, 90, 128)):
        det.add_template_pyramid("p%d" % (k % 3), pyr)
    gray = synth.scene_copies(1)[0]
    img = bgr_of(gray, 5)
    ref = oracle_tuples(oracle.match16(img, [4, 8], 60, bank_of(det), nthreads=NTHREADS))
    say("oracle matches", len(ref))
    assert len(ref) > 0
    shapes = [(16, 2, 1), (16, 2, 4), (8, 0, 1), (4, 1, 2), (32, 3, 1), (32, 3, 8)]
    for (ppl, gather, split) in shapes:
        det.set_tuning(positions_per_lane=ppl, gather=gather, split=split, waves_per_block=min(split, 4) if split > 1 else 4)
        for cmode in (0, 2):
            det.set_coarse_mode(cmode)
            got = tuples(det.match(img, 60))
            say("ppl", ppl, "gather", gather, "split", split, "coarse mode", cmode, "path", det.last_coarse_path())
            assert got == ref
    det.set_coarse_mode(1)
    det.set_tuning()
    dimgs = [det.upload_image(img), det.upload_image(bgr_of(synth.scene_copies(2)[0], 6))]
    refs = [ref, oracle_tuples(oracle.match16(bgr_of(synth.scene_copies(2)[0], 6), [4, 8], 60, bank_of(det),
                                              nthreads=NTHREADS))]
    for rep in range(3):  # eager, captured, replayed
        for k, d in enumerate(dimgs):
            assert tuples(det.match_image(d, 60)) == refs[k], (rep, k)
    got = det.match_images(dimgs, 60, slots=2)
    assert [tuples(g) for g in got] == refs


def test_match16_default_pyramid_and_mask():
    """Detector16() keeps the variant's T = {5, 8}: an 80-multiple crop, with a mask."""
    det = sbm.Detector16()
    assert [det.getT(0), det.getT(1)] == [5, 8]
    assert det.getModalities() == dict(weak_threshold=10.0, num_features=63, strong_threshold=55.0)
    img = train_image("case1_train.png")
    for a in range(0, 360, 60):
        det.addTemplate(oracle.shape_transform(img, a, 1.0), "t")
    scene = crop(scene_bgr(1), 80)
    rows, cols = scene.shape[:2]
    mask = np.zeros((rows, cols), np.uint8)
    mask[40:-40, 80:-80] = 255
    ref = oracle_tuples(oracle.match16(scene, [5, 8], 50, bank_of(det), mask=mask, nthreads=NTHREADS))
    got = tuples(det.match(scene, 50, masks=mask))
    say("T {5, 8}: matches", len(got))
    assert got == ref


def test_more_than_8191_features_per_level():
    g = np.random.default_rng(3)
    n0, n1 = 9000, 4500
    f0 = np.column_stack([g.integers(0, 90, n0), g.integers(0, 90, n0), g.integers(0, 16, n0)]).astype(np.int32)
    f1 = np.column_stack([g.integers(0, 45, n1), g.integers(0, 45, n1), g.integers(0, 16, n1)]).astype(np.int32)
    pyr = [dict(width=90, height=90, tl_x=0, tl_y=0, features=f0), dict(width=45, height=45, tl_x=0, tl_y=0, features=f1)]
    det = sbm.Detector16([4, 8])
    det.add_template_pyramid("big", pyr)
    img = crop(scene_bgr(0), 16)[:256, :384]
    ref = oracle_tuples(oracle.match16(img, [4, 8], 15, bank_of(det), nthreads=NTHREADS))
    got = tuples(det.match(img, 15))
    say("9000 features: matches", len(got))
    assert got == ref and len(ref) > 0
    det8 = sbm.Detector([4, 8])
    det8.add_template_pyramid("big", [dict(lv, features=np.column_stack([lv["features"][:, :2], lv["features"][:, 2] % 8]))
                                      for lv in pyr])
    with pytest.raises(sbm.SbmError):
        det8.match(np.ascontiguousarray(img[:, :, 0]), 15)  # 8 orientations stop at 8191


def test_variant_rules():
    det = sbm.Detector16([4, 8])
    img = train_image("case1_train.png")
    assert det.addTemplate(img, "a", None, 64) == 0
    with pytest.raises(sbm.SbmError):
        det.addTemplate_rotate("a", 0, 30.0, (100.0, 100.0))
    with pytest.raises(sbm.SbmError):
        det.match(np.zeros((64, 64), np.uint8), 50)  # single channel
    with pytest.raises(sbm.SbmError):
        det.addTemplate(np.zeros((64, 64), np.uint8), "a")
    assert det.orientationBins() == 16 and sbm.Detector().orientationBins() == 8
