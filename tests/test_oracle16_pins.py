"""Pins the 16-orientation oracle (oracle/sbm_oracle16.c) before it is trusted as the checker.

The variant is test/ori_16bit_experiment/line2Dup_16bit_ori.cpp.  Each check compares the oracle with
something it was not derived from: the tables the reference ships (the one compiled into the variant
and LUT16.txt beside it), closed forms (ramps whose gradient direction is known, the 8-orientation
oracle on the same ramps), direct formulations (window ORs with numpy shifts, placement sums over
response maps), and hand-computed micro cases for the driver and the feature selection.
"""
import json

import numpy as np
import pytest

import oracle
from helpers import golden


def lut16_tables():
    d = json.load(open(golden("similarity_lut16.json")))
    return np.array(d["compiled"], np.uint8), np.array(d["generated"], np.uint8)


def rnd(shape, seed):
    return np.random.default_rng(seed).integers(0, 256, shape, dtype=np.uint8)


def ramp(theta_deg, n=24, slope=7.0):
    yy, xx = np.mgrid[0:n, 0:n].astype(np.float64)
    th = np.deg2rad(theta_deg)
    g = np.clip(np.round(128 + slope * (np.cos(th) * (xx - n / 2) + np.sin(th) * (yy - n / 2))), 0, 255)
    g = g.astype(np.uint8)
    return np.stack([g, g, g], -1)


# ---------------------------------------------------------------------------------------------
# SIMILARITY_LUT[1024] (line2Dup_16bit_ori.cpp:575-609) and LUT16.txt (LUT_gen.cpp)
# ---------------------------------------------------------------------------------------------
def test_lut16_closed_form_equals_compiled_table():
    compiled, _ = lut16_tables()
    assert np.array_equal(oracle.similarity_lut16(), compiled)
    assert set(np.unique(compiled)) == {0, 1, 4}


def test_lut_gen_closed_form_equals_lut16_txt_which_the_variant_does_not_compile():
    compiled, generated = lut16_tables()
    assert np.array_equal(oracle.generated_lut16(), generated)
    assert generated.max() == 8 and not np.array_equal(compiled, generated)


# ---------------------------------------------------------------------------------------------
# quantization (hysteresisGradient, 16 bins: line2Dup_16bit_ori.cpp:216-296)
# ---------------------------------------------------------------------------------------------
@pytest.mark.parametrize("k", range(32))
def test_ramp_at_each_bin_centre_quantizes_to_its_bin(k):
    """A ramp rising along theta = k * 11.25 degrees has phase theta everywhere inside: bin k mod 16
    (32 buckets over 360 degrees folded by "& 15"); the 8-orientation oracle gives bin k/2 mod 8 on
    the even centres, whose 8-bin value is not a rounding tie."""
    img = ramp(k * 11.25)
    q, mag, _ = oracle.quantized_orientations16(img)
    inner = q[5:-5, 5:-5]
    assert (inner == (1 << (k % 16))).all()
    assert (q[0] == 0).all() and (q[-1] == 0).all() and (q[:, 0] == 0).all() and (q[:, -1] == 0).all()
    if k % 2 == 0:
        q8, mag8, _ = oracle.quantized_orientations(img, weak=10.0)
        assert (q8[5:-5, 5:-5] == (1 << ((k // 2) % 8))).all()
        assert np.array_equal(mag, mag8)


def test_weak_ramps_are_not_quantized():
    q, mag, _ = oracle.quantized_orientations16(ramp(33.75, slope=1.0))  # |grad| = 8: 64 <= 10^2
    assert (q == 0).all() and (mag <= 100).all()


def test_quantized16_values_and_vote():
    """Single bits only; a pixel carries bin b only where its magnitude passes and at least five of
    its 3x3 raw buckets are b (first maximum wins), checked with a direct count."""
    img = np.stack([rnd((40, 44), s) for s in (1, 2, 3)], -1)
    img = oracle.gaussian_blur7(img[:, :, 0])  # smooth, so that many votes pass
    img = np.stack([img, np.roll(img, 3, 1), np.roll(img, 5, 0)], -1)
    q, mag, ag = oracle.quantized_orientations16(img)
    assert set(np.unique(q)) <= {0} | {1 << b for b in range(16)}
    raw = np.array([[oracle.angle_bin16(a) for a in row] for row in ag], np.int64)
    raw[0, :] = raw[-1, :] = 0
    raw[:, 0] = raw[:, -1] = 0
    raw[1:-1, 1:-1] &= 15
    rows, cols = q.shape
    labelled = 0
    for r in range(1, rows - 1):
        for c in range(1, cols - 1):
            want = 0
            if mag[r, c] > 100.0:
                counts = np.bincount(raw[r - 1:r + 2, c - 1:c + 2].ravel(), minlength=16)
                if counts.max() >= 5:
                    want = 1 << int(np.argmax(counts))
            assert q[r, c] == want, (r, c)
            labelled += want != 0
    assert labelled > 100


def test_single_channel_image_is_refused():
    with pytest.raises(ValueError):
        oracle.quantized_orientations16(rnd((16, 16), 0))


# ---------------------------------------------------------------------------------------------
# spread (16-bit OR), responses (four nibble lookups, the fourth always nibble 0)
# ---------------------------------------------------------------------------------------------
@pytest.mark.parametrize("T", [1, 3, 4, 8])
def test_spread16_equals_shifted_or(T):
    rng = np.random.default_rng(T)
    q = ((1 << rng.integers(0, 16, (24, 32))) * (rng.random((24, 32)) < 0.2)).astype(np.uint16)
    p = np.pad(q, ((0, T), (0, T)))
    want = np.zeros_like(q)
    for dr in range(T):
        for dc in range(T):
            want |= p[dr:dr + 24, dc:dc + 32]
    assert np.array_equal(oracle.spread16(q, T), want)


def test_response16_follows_compiled_table():
    compiled, _ = lut16_tables()
    s = np.random.default_rng(7).integers(0, 1 << 16, (32, 32)).astype(np.uint16)
    s[0, :16] = 1 << np.arange(16)
    r = oracle.response_maps16(s)
    for ori in range(16):
        base = 64 * ori
        want = np.maximum.reduce([compiled[base + (s & 15)], compiled[base + 16 + ((s >> 4) & 15)],
                                  compiled[base + 32 + ((s >> 8) & 15)], np.full(s.shape, compiled[base + 48])])
        assert np.array_equal(r[ori], want)
    # line2Dup_16bit_ori.cpp:639 masks with 15 << 16, so bits 12..15 alone give no response
    assert (r[:, 0, 12:16] == 0).all()
    assert (r[:, 0, :12].max(axis=0) == 4).all()


# ---------------------------------------------------------------------------------------------
# similarity / similarityLocal over 16 orientations
# ---------------------------------------------------------------------------------------------
def flat_lm16(resp, T):
    parts = [oracle.linearize(resp[o], T).ravel() for o in range(16)]
    rows, cols = resp.shape[1:]
    return np.concatenate(parts + [np.zeros(oracle.lib().orc_lm_guard_bytes(rows, cols, T), np.uint8)])


def random_response16(rows, cols, seed):
    return np.random.default_rng(seed).choice(np.array([0, 1, 4], np.uint8), size=(16, rows, cols))


@pytest.mark.parametrize("T,n", [(4, 40), (8, 70), (5, 63)])
def test_similarity16_equals_direct_placement_sum(T, n):
    rows, cols = 16 * T, 12 * T
    resp = random_response16(rows, cols, T)
    lm = flat_lm16(resp, T)
    rng = np.random.default_rng(T + 1)
    width, height = 5 * T - 1, 4 * T + 2
    feats = np.column_stack([rng.integers(0, width + 1, n), rng.integers(0, height + 1, n),
                             rng.integers(0, 16, n)]).astype(np.int32)
    got = oracle.similarity16(lm, rows, cols, T, feats, width, height)
    W, H = cols // T, rows // T
    wf, hf = (width - 1) // T + 1, (height - 1) // T + 1
    P = (H - hf) * W + (W - wf) + 1
    checked = 0
    for y in range(H - hf + 1):
        for x in range(W - wf + 1):
            if any(y * T + f[1] >= rows or x * T + f[0] >= cols for f in feats):
                continue
            assert got[y, x] == sum(int(resp[f[2], y * T + f[1], x * T + f[0]]) for f in feats), (y, x)
            checked += 1
    assert checked > (H - hf) * (W - wf) // 2
    assert (got.ravel()[P:] == 0).all()


@pytest.mark.parametrize("T,n", [(4, 50), (8, 90)])
def test_similarity_local16_equals_direct_window_sum(T, n):
    rows, cols = 40 * T, 40 * T
    resp = random_response16(rows, cols, T + 30)
    lm = flat_lm16(resp, T)
    rng = np.random.default_rng(T)
    feats = np.column_stack([rng.integers(0, 6 * T, n), rng.integers(0, 6 * T, n), rng.integers(0, 16, n)])
    feats = feats.astype(np.int32)
    for cx, cy in [(8 * T, 8 * T), (13 * T + 3, 17 * T + 1)]:
        got = oracle.similarity_local16(lm, rows, cols, T, feats, cx, cy)
        ox, oy = (cx // T - 8) * T, (cy // T - 8) * T
        for r in range(16):
            for c in range(16):
                want = sum(int(resp[f[2], f[1] + oy + r * T, f[0] + ox + c * T]) for f in feats)
                assert got[r, c] == want


# ---------------------------------------------------------------------------------------------
# matchClass driver micro case (line2Dup_16bit_ori.cpp:1140-1299)
# ---------------------------------------------------------------------------------------------
def test_driver16_coarse_to_refined_position_and_score():
    rows, cols, T = 256, 256, [4, 8]
    lab = 13  # an orientation past the first eight
    f0 = np.array([[0, 0, lab]], np.int32)
    bank = {"c": [[dict(width=4, height=4, tl_x=0, tl_y=0, features=f0),
                   dict(width=2, height=2, tl_x=0, tl_y=0, features=f0)]]}
    lms = []
    for lvl, (t, entries) in enumerate(zip(T, [[(92, 100, 4), (88, 96, 1)], [(40, 48, 4)]])):
        resp = np.zeros((16, rows >> lvl, cols >> lvl), np.uint8)
        for (y, x, v) in entries:
            resp[lab, y, x] = v
        lms.append(flat_lm16(resp, t))
    res = oracle.match_lm16(lms, rows, cols, T, 90.0, bank)
    assert res == [((103 // 4 - 8 + 8) * 4 + 1, (87 // 4 - 8 + 10) * 4 + 1, 100.0, "c", 0)]
    assert oracle.match_lm16(lms, rows, cols, T, 100.0, bank) == []  # coarse keeps '>' only


# ---------------------------------------------------------------------------------------------
# training (extractTemplate / selectScatteredFeatures of the variant, :167-210, :421-498)
# ---------------------------------------------------------------------------------------------
def peaks_image(points, rows=64, cols=64):
    """Isolated magnitude peaks (score, label) at the given positions; everything else 0."""
    mag = np.zeros((rows, cols), np.float32)
    ang = np.zeros((rows, cols), np.uint16)
    for (y, x, s, lab) in points:
        mag[y, x] = s
        ang[y, x] = 1 << lab
    return mag, ang


def test_selection_stops_at_exactly_num_features():
    # six candidates, num_features 2: distance 6 // 2 + 1 = 4.  B lies 3 px from A (outside the 5x5
    # NMS window, inside the distance), so the first pass takes A then C and stops there.
    pts = [(10, 10, 9000.0, 1), (10, 13, 8000.0, 2), (30, 30, 7000.0, 3), (40, 40, 6000.0, 14),
           (50, 50, 5000.0, 5), (20, 50, 4000.0, 15)]
    mag, ang = peaks_image(pts)
    got = oracle.extract_template16(mag, ang, None, 55.0, 2)
    assert got.tolist() == [[10, 10, 1], [30, 30, 3]]
    got = oracle.extract_template16(mag, ang, None, 55.0, 5)  # distance 2: B kept in the first pass
    assert got.tolist() == [[10, 10, 1], [13, 10, 2], [30, 30, 3], [40, 40, 14], [50, 50, 5]]


def test_selection_relaxes_distance_until_full():
    # four candidates within a 4x4 grid step of 3 px, num_features 4: distance 2 takes all on pass one
    pts = [(20, 20, 900.0 * 9, 0), (20, 23, 800.0 * 9, 4), (23, 20, 700.0 * 9, 8), (23, 23, 600.0 * 9, 12)]
    mag, ang = peaks_image(pts)
    got = oracle.extract_template16(mag, ang, None, 55.0, 4)
    assert len(got) == 4 and sorted(got[:, 2].tolist()) == [0, 4, 8, 12]


def test_too_few_candidates_refuses_the_template():
    pts = [(10, 10, 9000.0, 1), (30, 30, 7000.0, 3), (50, 50, 5000.0, 5)]
    mag, ang = peaks_image(pts)
    assert oracle.extract_template16(mag, ang, None, 55.0, 4) is None  # no exhaustive mode
    assert oracle.extract_template16(mag, ang, None, 55.0, 3) is not None
    assert oracle.extract_template16(mag, ang, None, 95.0, 1) is None  # 9000 <= 95^2


def test_add_template16_counts_per_level():
    from shape_based_matching_amd import synth

    _, templ, _, _ = synth.scene_plain(0)
    img = np.stack([templ, np.roll(templ, 1, 1), templ], -1)
    pyr = oracle.add_template16(img, 30, num_levels=2)
    assert pyr is not None
    assert [len(lv["features"]) for lv in pyr] == [30, 15]  # num_features halves per level
    assert oracle.add_template16(img, 64, num_levels=2) is None  # 35 candidates at level 0
    for lv in pyr:
        assert lv["features"][:, 2].min() >= 0 and lv["features"][:, 2].max() <= 15
    assert oracle.add_template16(np.full((64, 64, 3), 90, np.uint8), 16) is None
