"""Parity of the HIP path (through the C ABI) with the CPU oracle, stage by stage and end to end.

Integer outputs must be bit-identical; the one float step (phase angle, and the magnitude, which is
an exact integer) must be bitwise identical too.
"""
import os

import numpy as np
import pytest

import oracle
import shape_based_matching_amd as sbm
from helpers import bank_of, load_fixture_detector, oracle_tuples, scene_bgr, to_gray, tuples
from shape_based_matching_amd import synth

pytestmark = pytest.mark.gpu
NTHREADS = max(1, min(16, os.cpu_count() or 1))


def say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def c0():
    img, templ, _, _ = synth.scene_plain(0)
    return img, templ


def oracle_levels(img, T, mask=None, weak=30.0):
    """Per-level masked quantized orientations from the oracle."""
    cur, m, out = img, mask, []
    for lvl in range(len(T)):
        if lvl > 0:
            cur = oracle.pyr_down(cur)
            m = None if m is None else oracle.mask_down(m)
        q, _, _ = oracle.quantized_orientations(cur, weak)
        if m is not None:
            q = q.copy()
            q[m == 0] = 0
        out.append((cur, q))
    return out


def test_bin_table_exhaustive():
    say("bin table [-1020, 1020]^2")
    det = sbm.Detector()
    gpu = det.bin_table(-1020, 1020)
    cpu = oracle.bin_table(-1020, 1020)
    bad = np.argwhere(gpu != cpu)
    assert len(bad) == 0, f"{len(bad)} (dx,dy) pairs differ, first {bad[:5] - 1020}"
    with pytest.raises(sbm.SbmError):
        det.bin_table(-1021, 1020)  # outside the Sobel range of 8-bit images


def test_phase_quotient_equals_ieee_division_on_its_domain():
    """The gradient kernels take fastAtan2's quotient num / (den + DBL_EPSILON) from one reciprocal and one
    correction step; over every integer pair 0 <= num <= den <= 1020 it must equal the IEEE quotient bit for bit."""
    assert sbm.Detector().division_check() == 0


def _images():
    img0, _, _, _ = synth.scene_plain(0)
    return {"synthetic640": img0, "case0_gray": to_gray(scene_bgr(0)), "case1_bgr": scene_bgr(1),
            "case2_bgr": scene_bgr(2), "case2_gray": to_gray(scene_bgr(2)),
            "wide672": synth.scene_plain(3, rows=480, cols=672)[0]}


@pytest.mark.parametrize("name", ["synthetic640", "case0_gray", "case1_bgr", "case2_bgr"])
def test_gradient_matches_oracle(name):
    img = _images()[name]
    say("gradient", name, img.shape)
    det = sbm.Detector()
    q, mag, ang = det.gradient(img)
    oq, omag, oang = oracle.quantized_orientations(img)
    assert np.array_equal(mag.view(np.uint32), omag.view(np.uint32))
    assert np.array_equal(ang.view(np.uint32), oang.view(np.uint32))
    assert np.array_equal(q, oq)


@pytest.mark.parametrize("name,use_mask,T", [("synthetic640", False, [4, 8]), ("synthetic640", True, [4, 8]),
                                             ("case1_bgr", False, [4, 8]), ("case2_gray", False, [4, 8]),
                                             ("synthetic640", False, [2, 4, 8]), ("synthetic640", False, [5]),
                                             ("synthetic640", False, [16]), ("wide672", False, [4, 8])])
def test_levels_match_oracle(name, use_mask, T):
    img = _images()[name]
    rows, cols = img.shape[:2]
    mask = None
    if use_mask:
        yy, xx = np.mgrid[0:rows, 0:cols]
        mask = (((xx - cols / 2) ** 2 + (yy - rows / 2) ** 2) < (0.4 * rows) ** 2).astype(np.uint8) * 255
    say("levels", name, img.shape, "mask" if use_mask else "", T)
    det = sbm.Detector(63, T)
    det.prepare_image(img, mask)
    lms = oracle.lm_pyramid(img, T, mask=mask)
    for lvl, (lvl_img, q) in enumerate(oracle_levels(img, T, mask)):
        assert np.array_equal(det.read_level_image(lvl), lvl_img), f"level {lvl} image"
        assert np.array_equal(det.read_quantized(lvl), q), f"level {lvl} quantized"
        r, c = q.shape
        assert np.array_equal(det.read_linear_memories(lvl), lms[lvl][:8 * r * c]), f"level {lvl} linear memories"


def _pack_nibbles(lm_block):
    b = lm_block.reshape(-1, 2).astype(np.uint16)
    return (b[:, 0] | (b[:, 1] << 4)).astype(np.uint8)


def test_nibble_level_equals_packed_linear_memories():
    """The coarse level's two-positions-per-byte copy (gather 3) equals the byte linear memories packed,
    guard included and zero, also for an image prepared after a larger one (stale guard) and for a
    width whose LM rows are not a multiple of four columns (separate packing kernel)."""
    det = sbm.Detector(63, [4, 8])
    det.set_tuning(32, 4, 2, 3, 1)
    for rows, cols, seed in ((1024, 1024, 1), (480, 640, 2), (480, 672, 3), (480, 640, 4)):
        img = synth.scene_plain(seed, rows=rows, cols=cols)[0]
        say("nibbles", rows, cols)
        det.prepare_image(img)
        lm = det.read_linear_memories(1)
        nib = det.read_nibbles(1)
        W, H = (cols // 2) // 8, (rows // 2) // 8
        guard = ((W * H + 17 * W + 4096) + 15) & ~15
        block = np.concatenate([lm, np.zeros(guard, np.uint8)])
        packed = _pack_nibbles(block[:len(block) // 32 * 32])
        assert np.array_equal(nib[:len(packed)], packed)
        if W % 4 == 0:  # the banded spread kernel writes the nibbles and clears the whole guard
            assert not nib[len(packed):].any()


def _wrapping_templates(det, cid):
    """Template ids whose lowest-level features reach x == width or y == height with the size
    divisible by T (the reads that run into the next phase row)."""
    L = det.pyramidLevels()
    T = det.getT(L - 1)
    out = []
    for i in range(det.numTemplates(cid)):
        t = det.getTemplates(cid, i)[L - 1]
        f = t.features
        if len(f) and ((t.width % T == 0 and (f[:, 0] == t.width).any()) or
                       (t.height % T == 0 and (f[:, 1] == t.height).any())):
            out.append(i)
    return out


@pytest.mark.parametrize("case", [0, 1, 2])
def test_similarity_maps_match_oracle(case):
    bgr = scene_bgr(case)
    img = to_gray(bgr)
    det = load_fixture_detector(case)
    cid = det.classIds()[0]
    det.prepare_image(img)
    T = [4, 8]
    lms = oracle.lm_pyramid(img, T)
    rows, cols = img.shape
    r1, c1 = rows // 2, cols // 2
    n = det.numTemplates(cid)
    ids = sorted(set([0, 1, n // 3, n // 2, n - 1] + _wrapping_templates(det, cid)[:6]))
    say("similarity maps case", case, "templates", ids)
    for tid in ids:
        pyr = det.getTemplates(cid, tid)
        t1 = pyr[1]
        ref = oracle.similarity(lms[1], r1, c1, 8, t1.features, t1.width, t1.height)
        got = det.similarity_map(cid, tid)
        assert np.array_equal(got, ref), f"template {tid}"
        t0 = pyr[0]
        for (cx, cy) in [(8 * 4, 8 * 4), (cols // 3, rows // 2), (cols - t0.width - 32, rows - t0.height - 32)]:
            cx, cy = max(cx, 32), max(cy, 32)
            refl = oracle.similarity_local(lms[0], rows, cols, 4, t0.features, cx, cy)
            gotl = det.similarity_local(cid, tid, 0, cx, cy)
            assert np.array_equal(gotl, refl), f"template {tid} local ({cx},{cy})"


@pytest.mark.parametrize("case,color", [(0, False), (0, True), (1, True), (2, True), (2, False)])
@pytest.mark.parametrize("mode", [sbm.MODE_PLAIN, sbm.MODE_BATCHED])
def test_fixture_match_lists_equal_oracle(case, color, mode):
    bgr = scene_bgr(case)
    img = bgr if color else to_gray(bgr)
    det = load_fixture_detector(case)
    say("match fixture", case, "color" if color else "gray", "mode", mode, img.shape)
    got = tuples(det.match(img, 90, mode=mode))
    ref = oracle_tuples(oracle.match(img, [4, 8], 90, bank_of(det), nthreads=NTHREADS))
    assert got == ref
    say("  matches", len(got))


def test_training_matches_oracle(c0):
    _, templ = c0
    det = sbm.Detector(128, [4, 8])
    tid = det.addTemplate(templ, "poly", None, 128)
    assert tid == 0
    ref = oracle.add_template(templ, 128, 2)
    got = det.getTemplates("poly", 0)
    for lvl in range(2):
        g, r = got[lvl], ref[lvl]
        assert (g.width, g.height, g.tl_x, g.tl_y) == (r["width"], r["height"], r["tl_x"], r["tl_y"])
        assert np.array_equal(g.features, r["features"])
        assert np.array_equal(g.theta.view(np.uint32), np.asarray(r["theta"], np.float32).view(np.uint32))
    # masked training, colour input
    bgr = np.ascontiguousarray(np.stack([templ, templ // 2 + 40, 255 - templ], axis=2))
    mask = np.zeros(templ.shape, np.uint8)
    mask[10:150, 12:150] = 255
    tid = det.addTemplate(bgr, "poly", mask, 64)
    ref = oracle.add_template(bgr, 64, 2, mask=mask)
    got = det.getTemplates("poly", tid)
    for lvl in range(2):
        assert np.array_equal(got[lvl].features, ref[lvl]["features"])


def test_plain_scene_end_to_end(c0):
    img, templ = c0
    det = sbm.Detector(128, [4, 8])
    assert det.addTemplate(templ, "poly", None, 128) == 0
    base = oracle.add_template(templ, 128, 2)
    for ang in range(1, 360):
        det.addTemplate_rotate("poly", 0, float(ang), (80.0, 80.0))
    assert det.numTemplates() == 360
    for ang in (1, 45, 90, 181, 359):
        ref = oracle.rotate_template(base, float(ang), (80.0, 80.0))
        got = det.getTemplates("poly", ang)
        for lvl in range(2):
            assert np.array_equal(got[lvl].features, ref[lvl]["features"])
            assert (got[lvl].width, got[lvl].height, got[lvl].tl_x, got[lvl].tl_y) == \
                (ref[lvl]["width"], ref[lvl]["height"], ref[lvl]["tl_x"], ref[lvl]["tl_y"])
    ref = oracle_tuples(oracle.match(img, [4, 8], 90, bank_of(det), nthreads=NTHREADS))
    for mode in (sbm.MODE_PLAIN, sbm.MODE_BATCHED):
        say("640x480 x 360 templates, mode", mode)
        assert tuples(det.match(img, 90, mode=mode)) == ref
    # lower threshold: many more candidates through refinement
    ref = oracle_tuples(oracle.match(img, [4, 8], 60, bank_of(det), nthreads=NTHREADS))
    assert tuples(det.match(img, 60, mode=sbm.MODE_BATCHED)) == ref
    say("  matches at 60:", len(ref))


def test_copies_scene_batched_equals_plain_and_oracle():
    img, templ, placed = synth.scene_copies(1)
    _, _, poly, _ = synth.scene_plain(0)
    det = sbm.Detector(256, [4, 8])
    for ang in range(360):
        det.add_template_pyramid("poly", synth.bank_template(poly, float(ang), 1.0, 60, 256))
    say("1024x1024 x 360 templates x 256 features")
    plain = tuples(det.match(img, 80, mode=sbm.MODE_PLAIN))
    batched = tuples(det.match(img, 80, mode=sbm.MODE_BATCHED))
    assert plain == batched
    ref = oracle_tuples(oracle.match(img, [4, 8], 80, bank_of(det), nthreads=NTHREADS))
    assert batched == ref
    say("  matches:", len(ref))
    assert len(ref) > 0


@pytest.mark.parametrize("ppl,wpb,bpc,gather,split", [
    (4, 1, 1, 0, 1), (4, 1, 1, 1, 1), (4, 1, 1, 2, 1), (8, 2, 4, 0, 1), (8, 2, 4, 1, 1), (8, 2, 4, 2, 1),
    (16, 4, 8, 0, 1), (16, 4, 8, 1, 1), (16, 4, 8, 2, 1), (16, 1, 32, 2, 1), (32, 4, 8, 3, 1), (32, 1, 16, 3, 1),
    (8, 2, 4, 0, 2), (16, 2, 4, 2, 2), (16, 4, 8, 1, 4), (32, 4, 8, 3, 4), (32, 2, 2, 3, 2), (32, 4, 2, 3, 8),
    (16, 4, 4, 2, 8), (4, 4, 2, 0, 8)])
def test_tile_shapes_agree(c0, ppl, wpb, bpc, gather, split):
    img, templ = c0
    det = load_fixture_detector(1)
    bgr = scene_bgr(1)
    det.set_tuning(ppl, wpb, bpc, gather, split)
    say("tuning", ppl, wpb, bpc, gather, split)
    got = tuples(det.match(bgr, 85, mode=sbm.MODE_BATCHED))
    plain_gather0 = det.match(bgr, 85, mode=sbm.MODE_PLAIN)
    det.set_tuning(*((32, 4, 8, 3) if gather == 2 else (16, 4, 8, 2)))
    assert got == tuples(plain_gather0) == tuples(det.match(bgr, 85, mode=sbm.MODE_PLAIN))


def test_autotune_keeps_results_and_picks_a_listed_shape():
    img, _, _ = synth.scene_copies(1)
    _, _, poly, _ = synth.scene_plain(0)
    det = sbm.Detector(256, [4, 8])
    for ang in range(0, 360, 4):
        det.add_template_pyramid("poly", synth.bank_template(poly, float(ang), 1.0, 60, 256))
    dimg = det.upload_image(img)
    before = tuples(det.match_image(dimg, 80))
    best, trace = det.autotune(dimg, 80, iters=1)
    gather = [r for r in trace if r["stage"] == "gather"]
    refine = [r for r in trace if r["stage"] == "refine"]
    say("autotune best", best, "of", len(gather), "gather and", len(refine), "refinement shapes")
    assert len(gather) == 240 and len(refine) == 140 and len(trace) == 380
    assert best in [(r["positions_per_lane"], r["waves_per_block"], r["blocks_per_cu"], r["gather"], r["split"])
                    for r in gather]
    assert det.get_tuning() == best
    keys = ("refine_waves", "refine_group", "refine_blocks_per_cu", "refine_layout")
    chosen = tuple(det.get_tuning_param(k) for k in keys)
    assert chosen in [tuple(r[k] for k in keys) for r in refine]
    assert {r["refine_layout"] for r in refine} == {0, 1, 2, 3, 4}
    assert tuples(det.match_image(dimg, 80)) == before == tuples(det.match(img, 80, mode=sbm.MODE_PLAIN))


def test_batch_graphs_equal_single_matches():
    _, _, poly, _ = synth.scene_plain(0)
    imgs = [synth.scene_sized(i, r, c, poly, n_shapes=12) for i, (r, c) in
            enumerate([(480, 640), (960, 1280), (480, 640), (1936, 2592), (960, 1280), (480, 640), (480, 640)])]
    det = sbm.Detector(128, [4, 8])
    for ang in range(0, 360, 3):
        det.add_template_pyramid("poly", synth.bank_template(poly, float(ang), 1.0, 60, 128))
    dimgs = [det.upload_image(im) for im in imgs]
    single = [tuples(det.match_image(d, 75)) for d in dimgs]
    assert any(len(s) for s in single)
    for slots in (1, 3):
        say("batch of", len(imgs), "mixed sizes, slots", slots)
        got = [tuples(m) for m in det.match_images(dimgs, 75, slots=slots)]
        assert got == single
    # again (graphs replayed, buffers resident), then one image against the oracle
    assert [tuples(m) for m in det.match_images(dimgs[::-1], 75, slots=2)] == single[::-1]
    ref = oracle_tuples(oracle.match(imgs[1], [4, 8], 75, bank_of(det), nthreads=NTHREADS))
    assert single[1] == ref


def test_error_paths():
    det = load_fixture_detector(2)
    img = np.zeros((1000, 1000), np.uint8)  # level 1 is 500x500, not divisible by T=8
    with pytest.raises(sbm.SbmError) as e:
        det.match(img, 90)
    assert e.value.code == -2
    det.add_template_pyramid("big", [dict(width=10, height=10, tl_x=0, tl_y=0,
                                          features=np.zeros((8192, 3), np.int32)),
                                     dict(width=5, height=5, tl_x=0, tl_y=0, features=np.zeros((10, 3), np.int32))])
    with pytest.raises(sbm.SbmError) as e:
        det.match(np.zeros((64, 64), np.uint8), 90)
    assert e.value.code == -3


def test_repeated_single_matches_replay_graphs_exactly():
    """A repeated single-image call is captured into a graph on its second occurrence and replayed
    afterwards; every replay, with eager work of other sizes and modes in between, equals the first
    (eager) result."""
    _, _, poly, _ = synth.scene_plain(0)
    det = sbm.Detector(128, [4, 8])
    for ang in range(0, 360, 6):
        det.add_template_pyramid("poly", synth.bank_template(poly, float(ang), 1.0, 60, 128))
    a = synth.scene_sized(7, 960, 1280, poly, n_shapes=12)
    b = synth.scene_sized(8, 480, 640, poly, n_shapes=12)
    da, db = det.upload_image(a), det.upload_image(b)
    ref_a = tuples(det.match_image(da, 75))
    ref_b = tuples(det.match(b, 75, mode=sbm.MODE_PLAIN))
    assert ref_a
    for k in range(6):
        say("replay", k)
        assert tuples(det.match_image(da, 75)) == ref_a
        if k % 2:
            assert tuples(det.match_image(db, 75)) == ref_b
        else:
            assert tuples(det.match(b, 75, mode=sbm.MODE_PLAIN)) == ref_b
        det.synchronize()
    assert tuples(det.match(a, 75)) == ref_a  # host input through the same graph


def test_strided_host_input_through_graph_replays():
    """A host image with row padding (stride > cols * channels), with a mask, called repeatedly: the
    eager call, the captured graph and its replays all equal the contiguous image's list."""
    import ctypes
    from shape_based_matching_amd._lib import check, lib
    _, _, poly, _ = synth.scene_plain(0)
    det = sbm.Detector(128, [4, 8])
    for ang in range(0, 360, 10):
        det.add_template_pyramid("poly", synth.bank_template(poly, float(ang), 1.0, 60, 128))
    img = synth.scene_sized(9, 480, 640, poly, n_shapes=12)
    bgr = np.ascontiguousarray(np.dstack([img, img[:, ::-1], 255 - img]))
    mask = np.zeros(img.shape, np.uint8)
    mask[20:-30, 30:-10] = 255
    for src in (img, bgr):
        ch = 1 if src.ndim == 2 else 3
        ref = tuples(det.match(src, 70, masks=mask))
        assert ref
        pad = np.full((src.shape[0], src.shape[1] * ch + 37), 7, np.uint8)
        pad[:, :src.shape[1] * ch] = src.reshape(src.shape[0], -1)
        ids = (ctypes.c_char_p * 1)()
        for k in range(4):
            say("strided host input, channels", ch, "call", k)
            out = ctypes.c_void_p()
            check(lib().sbm_match(det._h, pad.ctypes.data, src.shape[0], src.shape[1], pad.shape[1], ch, 70.0, ids, 0,
                                  mask.ctypes.data, sbm.MODE_BATCHED, ctypes.byref(out)))
            xyct, sim, names = det._collect(out)
            got = [(int(r[0]), int(r[1]), np.float32(s), names[r[2]], int(r[3])) for r, s in zip(xyct, sim)]
            assert got == ref
