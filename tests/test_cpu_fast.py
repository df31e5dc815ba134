"""The vectorised, OpenMP CPU baseline (oracle/sbm_cpu_fast.c) against the scalar oracle.

The baseline is timed by bench.py beside the GPU, so it must compute exactly what the scalar oracle
computes: the exhaustive (dx, dy) -> bin table, every level's linear memories (gray, colour, masked, other
T), and full match lists on the reference's fixture banks and on synthetic workloads, at one and several
threads.
"""
import numpy as np
import pytest

import oracle
from helpers import bank_of, load_fixture_detector, scene_bgr, to_gray
from shape_based_matching_amd import synth


def crop(img, m=16):
    r, c = img.shape[:2]
    return np.ascontiguousarray(img[: r - r % m, : c - c % m])


def test_isa_is_reported():
    assert oracle.fast_isa() in ("AVX-512", "AVX2", "x86-64")


def test_bin_table_exhaustive():
    assert np.array_equal(oracle.fast_bin_table(), oracle.bin_table())


@pytest.mark.parametrize("T", [[4, 8], [5, 8], [2, 4, 8], [16]])
def test_linear_memories_equal(T):
    _, _, poly, _ = synth.scene_plain(0)
    rows = 400 if 5 in T else 480
    img = synth.scene_sized(9, rows, 640, poly, n_shapes=25)
    yy, xx = np.mgrid[0:rows, 0:640]
    mask = ((xx - 300) ** 2 + (yy - 200) ** 2 < 180 ** 2).astype(np.uint8) * 255
    for m in (None, mask):
        a = oracle.lm_pyramid(img, T, mask=m)
        b = oracle.lm_pyramid_fast(img, T, mask=m)
        assert all(np.array_equal(x, y) for x, y in zip(a, b)), (T, m is None)


def test_linear_memories_colour_and_noise():
    g = np.random.default_rng(4)
    noise = g.integers(0, 256, (256, 320), dtype=np.uint8)
    bgr = np.ascontiguousarray(g.integers(0, 256, (256, 320, 3), dtype=np.uint8))
    for img in (noise, bgr):
        a = oracle.lm_pyramid(img, [4, 8])
        b = oracle.lm_pyramid_fast(img, [4, 8])
        assert all(np.array_equal(x, y) for x, y in zip(a, b))


@pytest.mark.parametrize("case", [0, 1, 2])
def test_fixture_banks_gray_and_colour(case):
    det = load_fixture_detector(case)
    bank = bank_of(det)
    fb = oracle.FlatBank(bank, 2)
    bgr = crop(scene_bgr(case))
    for img in (to_gray(bgr), bgr):
        ref = oracle.match(img, [4, 8], 75, bank, nthreads=4)
        for nt in (1, 4):
            got, _ = oracle.match_fast(img, [4, 8], 75, fb, nthreads=nt)
            assert got == ref, (case, img.ndim, nt)


def test_synthetic_bank_masks_and_thresholds():
    img = synth.scene_copies(1)[0]
    _, _, poly, _ = synth.scene_plain(0)
    bank = {"a": [synth.bank_template(poly, float(a), 1.0, 60, 256) for a in range(0, 360, 3)],
            "b": [synth.bank_template(poly, float(a), 0.9, 60, 40) for a in range(0, 360, 7)]}
    fb = oracle.FlatBank(bank, 2)
    mask = np.zeros(img.shape, np.uint8)
    mask[64:-96, 32:-160] = 255
    for thr, m in ((80, None), (60, None), (70, mask)):
        ref = oracle.match(img, [4, 8], thr, bank, mask=m, nthreads=4)
        got, times = oracle.match_fast(img, [4, 8], thr, fb, mask=m, nthreads=4)
        assert got == ref and len(ref) > 0, thr
        assert times[0] > 0 and times[1] > 0
    # a class selection follows the selected order, as orc_match does
    fsel = oracle.FlatBank(bank, 2, class_ids=["b"])
    assert oracle.match_fast(img, [4, 8], 60, fsel, nthreads=2)[0] == oracle.match(img, [4, 8], 60, bank, ["b"],
                                                                                   nthreads=2)


def test_limit_features_and_three_levels():
    g = np.random.default_rng(12)
    _, templ, poly, _ = synth.scene_plain(0)
    img = synth.scene_sized(13, 480, 640, poly, n_shapes=20)
    f = np.column_stack([np.full(8191, 20), np.full(8191, 20), np.arange(8191) % 8]).astype(np.int32)
    r = np.column_stack([g.integers(0, 40, 8191), g.integers(0, 40, 8191), g.integers(0, 8, 8191)]).astype(np.int32)
    bank = {"lim": [[dict(width=40, height=40, tl_x=0, tl_y=0, features=f),
                     dict(width=20, height=20, tl_x=0, tl_y=0, features=f)],
                    [dict(width=40, height=40, tl_x=0, tl_y=0, features=r),
                     dict(width=40, height=40, tl_x=0, tl_y=0, features=r)]]}
    ref = oracle.match(img, [4, 8], 20, bank, nthreads=4)
    assert oracle.match_fast(img, [4, 8], 20, oracle.FlatBank(bank, 2), nthreads=4)[0] == ref and len(ref) > 0
    bank3 = {"p": [[dict(width=w, height=w, tl_x=0, tl_y=0,
                         features=np.column_stack([g.integers(0, w, 60), g.integers(0, w, 60),
                                                   g.integers(0, 8, 60)]).astype(np.int32))
                    for w in (80, 40, 20)] for _ in range(12)]}
    ref = oracle.match(img, [2, 4, 8], 30, bank3, nthreads=4)
    assert oracle.match_fast(img, [2, 4, 8], 30, oracle.FlatBank(bank3, 3), nthreads=4)[0] == ref


def test_refusals_match_the_oracle():
    fb = oracle.FlatBank({"a": [synth.bank_template(synth.scene_plain(0)[2], 0.0, 1.0, 60, 64)]}, 2)
    with pytest.raises(ValueError):
        oracle.match_fast(np.zeros((470, 640), np.uint8), [4, 8], 80, fb)
