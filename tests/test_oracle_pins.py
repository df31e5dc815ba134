"""Pins the CPU oracle before it is trusted as the checker.

Each check compares the oracle with something it was not derived from: the reference's literal data
(SIMILARITY_LUT, the rotation-built fixture bank), an independent formulation (direct 2-D sums with
numpy padding modes instead of separable clamped loops, reshape/transpose instead of scatter loops),
closed forms (constant images, exact atan2 within the polynomial's error), or hand-computed micro
cases.  Each is chosen so that a plausible slip (wrong border mode, swapped dx/dy, off-by-one in the
over-read, '>' for '>=') fails.
"""
import math

import numpy as np
import pytest

import oracle
from helpers import bank_of, load_fixture_detector, lut_table

K7 = np.array([4, 14, 28, 36, 28, 14, 4], np.int64)
W5 = np.array([1, 4, 6, 4, 1], np.int64)


def rnd(shape, seed):
    return np.random.default_rng(seed).integers(0, 256, shape, dtype=np.uint8)


# ---------------------------------------------------------------------------------------------
# SIMILARITY_LUT (line2Dup.cpp:605-608)
# ---------------------------------------------------------------------------------------------
def test_lut_closed_form_equals_reference_table():
    assert np.array_equal(oracle.similarity_lut(), lut_table())


# ---------------------------------------------------------------------------------------------
# GaussianBlur 7x7 sigma 0 REPLICATE, Sobel REPLICATE, pyrDown REFLECT_101
# ---------------------------------------------------------------------------------------------
def blur_direct(img):
    k2 = np.outer(K7, K7)
    p = np.pad(img.astype(np.int64), 3, mode="edge")
    s = np.zeros(img.shape, np.int64)
    for i in range(7):
        for j in range(7):
            s += k2[i, j] * p[i:i + img.shape[0], j:j + img.shape[1]]
    return ((s + 8192) >> 14).astype(np.uint8)


@pytest.mark.parametrize("shape,seed", [((37, 53), 0), ((16, 16), 1), ((5, 9), 2), ((64, 48), 3)])
def test_blur_equals_direct_2d_sum(shape, seed):
    img = rnd(shape, seed)
    assert np.array_equal(oracle.gaussian_blur7(img), blur_direct(img))


def test_blur_constant_and_kernel_sum():
    assert K7.sum() ** 2 == 16384
    for v in (0, 1, 127, 255):
        img = np.full((20, 24), v, np.uint8)
        assert np.array_equal(oracle.gaussian_blur7(img), img)


@pytest.mark.parametrize("seed", [0, 1])
def test_sobel_equals_direct_correlation(seed):
    b = rnd((31, 45), seed)
    p = np.pad(b.astype(np.int64), 1, mode="edge")
    kx = np.array([[-1, 0, 1], [-2, 0, 2], [-1, 0, 1]])
    ky = kx.T
    dx = sum(kx[i, j] * p[i:i + b.shape[0], j:j + b.shape[1]] for i in range(3) for j in range(3))
    dy = sum(ky[i, j] * p[i:i + b.shape[0], j:j + b.shape[1]] for i in range(3) for j in range(3))
    gx, gy = oracle.sobel3(b)
    assert np.array_equal(gx, dx) and np.array_equal(gy, dy)
    # a left-dark / right-bright step has dx > 0, dy == 0 (orientation 0 degrees)
    step = np.zeros((8, 8), np.uint8)
    step[:, 4:] = 200
    gx, gy = oracle.sobel3(step)
    assert gx[4, 3] == 800 and gy[4, 3] == 0


@pytest.mark.parametrize("shape,ch,seed", [((32, 48), 1, 0), ((18, 22), 1, 1), ((16, 20), 3, 2)])
def test_pyr_down_equals_reflect101_direct(shape, ch, seed):
    img = rnd(shape + ((ch,) if ch > 1 else ()), seed)
    k2 = np.outer(W5, W5)
    pad = ((2, 2), (2, 2)) + (((0, 0),) if ch > 1 else ())
    p = np.pad(img.astype(np.int64), pad, mode="reflect")  # numpy 'reflect' is REFLECT_101
    s = np.zeros(img.shape, np.int64)
    for i in range(5):
        for j in range(5):
            s += k2[i, j] * p[i:i + img.shape[0], j:j + img.shape[1]]
    want = ((s[0::2, 0::2] + 128) >> 8).astype(np.uint8)[: shape[0] // 2, : shape[1] // 2]
    assert np.array_equal(oracle.pyr_down(img), want)


def test_mask_down_is_even_sample():
    m = rnd((20, 30), 4)
    assert np.array_equal(oracle.mask_down(m), m[0::2, 0::2])


# ---------------------------------------------------------------------------------------------
# phase (fastAtan2, degrees) and the 16 -> 8 bucket quantization
# ---------------------------------------------------------------------------------------------
def test_fast_atan2_special_values():
    assert oracle.fast_atan2_deg(0, 1) == 0.0
    assert oracle.fast_atan2_deg(1, 0) == 90.0
    assert oracle.fast_atan2_deg(0, -1) == 180.0
    assert oracle.fast_atan2_deg(-1, 0) == 270.0
    assert oracle.fast_atan2_deg(0, 0) == 0.0
    assert abs(oracle.fast_atan2_deg(1, 1) - 45.0) < 0.01
    assert abs(oracle.fast_atan2_deg(-1, -1) - 225.0) < 0.01


def test_fast_atan2_error_bound_and_bins():
    """Across the whole Sobel domain the polynomial stays within 0.01 degree of atan2, and its bins
    differ from exact-atan2 bins only on pairs whose exact angle sits within that error of a
    bucket boundary."""
    lo, hi = -1020, 1020
    table = oracle.bin_table(lo, hi)
    dy, dx = np.mgrid[lo:hi + 1, lo:hi + 1]
    exact = np.degrees(np.arctan2(dy, dx)) % 360.0
    q = exact * 16.0 / 360.0
    exact_bin = np.floor(q + 0.5).astype(np.int64) & 7
    diff = table != exact_bin
    frac = np.abs(q - np.floor(q) - 0.5)
    assert diff.sum() > 0  # the polynomial is not atan2
    assert (frac[diff] < 0.01 * 16 / 360 + 1e-6).all()
    assert diff.sum() < 2000
    # spot-check the float restatement against the error bound on a sample
    rng = np.random.default_rng(0)
    for _ in range(2000):
        x, y = (int(v) for v in rng.integers(lo, hi + 1, 2))
        a = oracle.fast_atan2_deg(y, x)
        e = math.degrees(math.atan2(y, x)) % 360.0
        d = abs(a - e)
        assert min(d, 360 - d) < 0.011


def test_angle_bin_rounds_half_to_even():
    step = np.float32(16.0 / 360.0)
    # 11.25 deg -> 0.5 bucket -> 0 (half to even); 33.75 -> 1.5 -> 2
    assert oracle.angle_bin(11.25) == int(np.rint(np.float32(11.25) * step))
    assert oracle.angle_bin(33.75) == 2
    assert oracle.angle_bin(360.0) == 16


# ---------------------------------------------------------------------------------------------
# hysteresisGradient (line2Dup.cpp:206-284)
# ---------------------------------------------------------------------------------------------
@pytest.mark.parametrize("kind,bit", [("vertical", 1), ("horizontal", 16), ("diag", 4), ("antidiag", 64)])
def test_quantized_orientation_of_straight_edges(kind, bit):
    n = 48
    yy, xx = np.mgrid[0:n, 0:n]
    if kind == "vertical":
        img = np.where(xx < n // 2, 40, 200)
    elif kind == "horizontal":
        img = np.where(yy < n // 2, 40, 200)
    elif kind == "diag":
        img = np.where(xx + yy < n, 40, 200)
    else:
        img = np.where(xx - yy < 0, 200, 40)
    q, mag, _ = oracle.quantized_orientations(img.astype(np.uint8))
    inner = q[8:-8, 8:-8]  # away from the replicated borders, which bend the gradient
    assert set(np.unique(inner)) <= {0, bit}
    assert (inner == bit).sum() > n // 2  # the edge is labelled
    assert (q[0] == 0).all() and (q[-1] == 0).all() and (q[:, 0] == 0).all() and (q[:, -1] == 0).all()
    assert (q[mag <= 900] == 0).all()  # weak threshold 30^2


def test_quantized_values_are_single_bits():
    q, _, _ = oracle.quantized_orientations(rnd((40, 40), 9))
    assert set(np.unique(q)) <= {0, 1, 2, 4, 8, 16, 32, 64, 128}


def test_flat_image_has_no_orientations():
    q, mag, _ = oracle.quantized_orientations(np.full((32, 32), 77, np.uint8))
    assert (q == 0).all() and (mag == 0).all()


# ---------------------------------------------------------------------------------------------
# spread, response maps, linearize (line2Dup.cpp:556-750)
# ---------------------------------------------------------------------------------------------
@pytest.mark.parametrize("T", [1, 2, 4, 8])
def test_spread_equals_shifted_or(T):
    rng = np.random.default_rng(T)
    q = (1 << rng.integers(0, 8, (24, 32))).astype(np.uint8) * (rng.random((24, 32)) < 0.2)
    q = q.astype(np.uint8)
    p = np.pad(q, ((0, T), (0, T)))
    want = np.zeros_like(q)
    for dr in range(T):
        for dc in range(T):
            want |= p[dr:dr + 24, dc:dc + 32]
    assert np.array_equal(oracle.spread(q, T), want)


def test_response_maps_follow_reference_lut():
    s = np.arange(256, dtype=np.uint8).reshape(16, 16)
    lut = lut_table()
    r = oracle.response_maps(s)
    for ori in range(8):
        want = np.maximum(lut[32 * ori + (s & 15)], lut[32 * ori + 16 + (s >> 4)])
        assert np.array_equal(r[ori], want)


@pytest.mark.parametrize("T", [1, 4, 8])
def test_linearize_is_phase_major_transpose(T):
    resp = rnd((8 * T, 4 * T), T)
    H, W = resp.shape[0] // T, resp.shape[1] // T
    want = resp.reshape(H, T, W, T).transpose(1, 3, 0, 2).reshape(T * T, H * W)
    assert np.array_equal(oracle.linearize(resp, T), want)


# ---------------------------------------------------------------------------------------------
# similarity / similarityLocal against the response maps directly
# ---------------------------------------------------------------------------------------------
def random_response(rows, cols, seed):
    rng = np.random.default_rng(seed)
    return rng.choice(np.array([0, 3, 4], np.uint8), size=(8, rows, cols))


def flat_lm(resp, T):
    parts = [oracle.linearize(resp[o], T).ravel() for o in range(8)]
    rows, cols = resp.shape[1:]
    guard = oracle.lib().orc_lm_guard_bytes(rows, cols, T)
    return np.concatenate(parts + [np.zeros(guard, np.uint8)])


@pytest.mark.parametrize("T,seed", [(4, 0), (8, 1), (2, 2)])
def test_similarity_equals_direct_placement_sum(T, seed):
    rows, cols = 16 * T, 12 * T
    resp = random_response(rows, cols, seed)
    lm = flat_lm(resp, T)
    rng = np.random.default_rng(seed + 10)
    width, height = 5 * T - 1, 4 * T + 2
    n = 70
    feats = np.column_stack([rng.integers(0, width + 1, n), rng.integers(0, height + 1, n),
                             rng.integers(0, 8, n)]).astype(np.int32)
    got = oracle.similarity(lm, rows, cols, T, feats, width, height)
    W, H = cols // T, rows // T
    wf, hf = (width - 1) // T + 1, (height - 1) // T + 1
    P = (H - hf) * W + (W - wf) + 1
    # positions where the template sits fully inside: the sum over features of the response at
    # (y*T + f.y, x*T + f.x)
    checked = 0
    for y in range(H - hf + 1):
        for x in range(W - wf + 1):
            if any(y * T + f[1] >= rows or x * T + f[0] >= cols for f in feats):
                continue  # an over-read position (covered by the tests below)
            want = sum(int(resp[f[2], y * T + f[1], x * T + f[0]]) for f in feats)
            assert got[y, x] == want, (y, x)
            checked += 1
    assert checked > (H - hf) * (W - wf) // 2
    assert (got.ravel()[P:] == 0).all()


def test_similarity_overread_into_next_phase_row():
    """A feature at y == height with height % T == 0 reads past its phase row on the last template row:
    the bytes come from the next phase row of the same orientation (line2Dup.cpp:798, 804-830)."""
    T, rows, cols = 4, 32, 32
    resp = random_response(rows, cols, 11)
    lm = flat_lm(resp, T)
    feats = np.array([[0, 8, 5]], np.int32)  # phase (0, 0), LM row offset 2*W
    got = oracle.similarity(lm, rows, cols, T, feats, 4, 8).ravel()
    W = cols // T
    P = (8 - 2) * W + (W - 1) + 1
    for j in range(P):
        y, x = divmod(j, W)
        if j + 2 * W < W * (rows // T):
            want = resp[5, (y + 2) * T, x * T]
        else:  # next phase row (grid (0, 1)), entry j + 2W - H*W
            k = j + 2 * W - W * (rows // T)
            want = resp[5, (k // W) * T, (k % W) * T + 1]
        assert got[j] == want, j


def test_similarity_wrapped_position_reads_next_lm_row():
    """A feature at (a*T, b*T) maps position (y, x) to LM row y+b, column x+a; past the right edge
    the read continues at row y+b+1, column x+a-W (the reference never filters these)."""
    T, rows, cols = 4, 32, 32
    resp = random_response(rows, cols, 5)
    lm = flat_lm(resp, T)
    a, b = 3, 1
    feats = np.array([[a * T, b * T, 6]], np.int32)
    width, height = a * T + 1, b * T + 1
    got = oracle.similarity(lm, rows, cols, T, feats, width, height).ravel()
    W, H = cols // T, rows // T
    wf, hf = (width - 1) // T + 1, (height - 1) // T + 1
    P = (H - hf) * W + (W - wf) + 1
    for j in range(P):
        y, x = divmod(j, W)
        yy, xx = (y + b, x + a) if x + a < W else (y + b + 1, x + a - W)
        assert got[j] == resp[6, yy * T, xx * T], j


@pytest.mark.parametrize("T", [4, 8])
def test_similarity_local_equals_direct_window_sum(T):
    rows, cols = 40 * T, 40 * T
    resp = random_response(rows, cols, T + 20)
    lm = flat_lm(resp, T)
    rng = np.random.default_rng(T)
    n = 90
    feats = np.column_stack([rng.integers(0, 6 * T, n), rng.integers(0, 6 * T, n), rng.integers(0, 8, n)]).astype(np.int32)
    for cx, cy in [(8 * T, 8 * T), (13 * T + 3, 17 * T + 1), (20 * T, 10 * T + T - 1)]:
        got = oracle.similarity_local(lm, rows, cols, T, feats, cx, cy)
        ox, oy = (cx // T - 8) * T, (cy // T - 8) * T
        for r in range(16):
            for c in range(16):
                want = sum(int(resp[f[2], f[1] + oy + r * T, f[0] + ox + c * T]) for f in feats)
                assert got[r, c] == want


# ---------------------------------------------------------------------------------------------
# matchClass driver micro cases (line2Dup.cpp:1131-1268)
# ---------------------------------------------------------------------------------------------
def one_feature_bank(n0=1):
    f0 = np.array([[0, 0, 2]] * n0, np.int32)
    f1 = np.array([[0, 0, 2]], np.int32)
    return {"c": [[dict(width=4, height=4, tl_x=0, tl_y=0, features=f0),
                   dict(width=2, height=2, tl_x=0, tl_y=0, features=f1)]]}


def lms_with(rows, cols, T, sets):
    out = []
    for lvl, (t, entries) in enumerate(zip(T, sets)):
        r, c = rows >> lvl, cols >> lvl
        resp = np.zeros((8, r, c), np.uint8)
        for (ori, y, x, v) in entries:
            resp[ori, y, x] = v
        out.append(flat_lm(resp, t))
    return out


def test_driver_coarse_to_refined_position_and_score():
    rows, cols, T = 256, 256, [4, 8]
    # level 1: response 4 at pixel (y, x) = (40, 48) for orientation 2 -> coarse cell (5, 6)
    # level 0: best response 4 at (81 + 4*2, 97 + 4*1) relative to the refined window
    lms = lms_with(rows, cols, T, [[(2, 92, 100, 4), (2, 88, 96, 3)], [(2, 40, 48, 4)]])
    res = oracle.match_lm(lms, rows, cols, T, 90.0, one_feature_bank())
    # coarse: x = 6*8 + 3 = 51, y = 5*8 + 3 = 43 -> refine centre (103, 87); window origin
    # ((103//4 - 8)*4, (87//4 - 8)*4) = (68, 52); best at column (100-68)/4 = 8, row (92-52)/4 = 10
    assert res == [((103 // 4 - 8 + 8) * 4 + 1, (87 // 4 - 8 + 10) * 4 + 1, 100.0, "c", 0)]


def test_driver_threshold_is_strict_coarse_inclusive_refine():
    rows, cols, T = 256, 256, [4, 8]
    lms = lms_with(rows, cols, T, [[(2, 92, 100, 3)], [(2, 40, 48, 4)]])
    # refine score = 3 * 100 / 4 = 75: kept at threshold 75 (>=), dropped just above
    assert len(oracle.match_lm(lms, rows, cols, T, 75.0, one_feature_bank())) == 1
    assert len(oracle.match_lm(lms, rows, cols, T, 75.0001, one_feature_bank())) == 0
    # coarse score 100 passes '>' 99.99 but not '>' 100
    lms2 = lms_with(rows, cols, T, [[(2, 92, 100, 4)], [(2, 40, 48, 4)]])
    assert len(oracle.match_lm(lms2, rows, cols, T, 99.99, one_feature_bank())) == 1
    assert len(oracle.match_lm(lms2, rows, cols, T, 100.0, one_feature_bank())) == 0


def test_driver_empty_refine_window_keeps_zero_at_nonpositive_threshold():
    rows, cols, T = 256, 256, [4, 8]
    lms = lms_with(rows, cols, T, [[], [(2, 40, 48, 4)]])
    assert oracle.match_lm(lms, rows, cols, T, 50.0, one_feature_bank()) == []
    res = oracle.match_lm(lms, rows, cols, T, 0.0, one_feature_bank())
    # coarse candidates: every cell with score > 0 -> only the one; best_c = best_r = -1
    assert res == [((103 // 4 - 9) * 4 + 1, (87 // 4 - 9) * 4 + 1, 0.0, "c", 0)]


def test_score_expression():
    assert oracle.score(3, 1) == 75.0
    assert oracle.score(1, 3) == np.float32(np.float32(100.0) / np.float32(12.0))
    assert math.isnan(oracle.score(0, 0))


# ---------------------------------------------------------------------------------------------
# canonical sort + unique, NMS (test.cpp:31-121)
# ---------------------------------------------------------------------------------------------
def test_sort_unique_canonical_order():
    m = [(10, 10, 90.0, "b", 2), (10, 10, 90.0, "a", 1), (5, 5, 95.0, "a", 7), (10, 10, 90.0, "a", 1),
         (4, 4, 90.0, "a", 1), (10, 10, 90.0, "a", 3)]
    got = oracle.sort_unique(m, ["a", "b"])
    # similarity desc, template asc, class asc, y asc, x asc; unique merges equal (x,y,sim,class)
    # the exact duplicate of template 1 merges; template 3's copy of (10, 10, 90, "a") is not adjacent
    # to template 1's (template 2 of class "b" sits between them), so it stays, as in the reference
    assert got == [(5, 5, 95.0, "a", 7), (4, 4, 90.0, "a", 1), (10, 10, 90.0, "a", 1), (10, 10, 90.0, "b", 2),
                   (10, 10, 90.0, "a", 3)]


def test_nms_hand_cases():
    boxes = [(0, 0, 10, 10), (0, 5, 10, 10), (1, 0, 10, 10), (0, 0, 10, 20), (50, 50, 5, 5)]
    scores = [0.9, 0.8, 0.95, 0.7, 0.0]
    # order by score: 2 (0.95), 0, 1, 3; box 4 is not > 0
    # 0 vs 2: inter 90, union 110 -> 0.818 > 0.5 suppressed; 1 vs 2: inter 45 / 155 -> kept;
    # 3 vs 2: inter 90 / 210 = 0.43 kept, vs 1: inter 100 / 200 = 0.5 -> kept (<= 0.5)
    assert oracle.nms_boxes(boxes, scores, 0.0, 0.5) == [2, 1, 3]
    # top_k truncates the sorted candidates before suppression (test.cpp:56-59): [2, 0] -> [2]
    assert oracle.nms_boxes(boxes, scores, 0.0, 0.5, top_k=2) == [2]
    assert oracle.nms_boxes(boxes, scores, 0.0, 0.5, top_k=3) == [2, 1]
    # equal scores keep input order (stable sort)
    assert oracle.nms_boxes([(0, 0, 4, 4), (100, 0, 4, 4)], [0.5, 0.5], 0.0, 0.5) == [0, 1]


# ---------------------------------------------------------------------------------------------
# training: addTemplate_rotate reproduces the reference's rotation-built bank
# ---------------------------------------------------------------------------------------------
def test_rotation_reproduces_fixture_bank_geometry():
    """test/case1 was trained by addTemplate + addTemplate_rotate(angle, centre (235, 235))
    (test.cpp:230-274).  Rotating template 0 by each angle must give every template's feature
    positions, width, height and top-left exactly."""
    det = load_fixture_detector(1)
    bank = bank_of(det)["test"]
    base = bank[0]
    for k in range(1, 361):
        rot = oracle.rotate_template(base, float(k), (235.0, 235.0))
        for lvl in range(2):
            a, b = rot[lvl], bank[k][lvl]
            assert np.array_equal(a["features"][:, :2], np.asarray(b["features"])[:, :2]), (k, lvl)
            assert (a["width"], a["height"], a["tl_x"], a["tl_y"]) == (b["width"], b["height"], b["tl_x"], b["tl_y"])


def test_rotation_labels_consistent_with_one_theta_per_feature():
    """The fixture stores no theta, but the rotated labels of all 360 templates must be explained by a
    single source angle per feature under label = int(theta' * 16 / 360 + 0.5) & 7."""
    det = load_fixture_detector(1)
    bank = bank_of(det)["test"]
    grid = np.arange(0, 360, 0.05, dtype=np.float32)
    ks = np.arange(1, 361, dtype=np.float32)
    for lvl in range(2):
        labels = np.stack([np.asarray(bank[k][lvl]["features"])[:, 2] for k in range(1, 361)])  # (360, n)
        for i in range(0, labels.shape[1], 5):
            th = grid[:, None] - ks[None, :]
            th = np.where(th < 0, th + np.float32(360), th)
            lab = (np.floor(th * np.float32(16) / np.float32(360) + np.float32(0.5)).astype(np.int64)) & 7
            ok = (lab == labels[:, i][None, :]).all(axis=1)
            assert ok.any(), (lvl, i)
