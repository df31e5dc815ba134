"""Batched coarse stage over the coarse level's nonzero responses (sbm_sparse.hip).

The nonzero-response path must return exactly the matches of the gather kernel and of the oracle:
on the reference's fixture banks, on synthetic banks over scenes whose coarse level is sparse and
dense, on an image whose coarse level needs two score-map segments per template, and through captured
graph replays.  The path taken is reported per match and follows the coarse mode and the density.
"""
import os

import numpy as np
import pytest

import oracle
import shape_based_matching_amd as sbm
from helpers import bank_of, load_fixture_detector, oracle_tuples, scene_bgr, tuples
from shape_based_matching_amd import synth

pytestmark = pytest.mark.gpu
NTHREADS = max(1, min(16, os.cpu_count() or 1))


def say(*a):
    print(*a, flush=True)


def crop16(img):
    r, c = img.shape[:2]
    return np.ascontiguousarray(img[: r - r % 16, : c - c % 16])


def run_modes(det, img, thr, class_ids=None):
    """Match lists under the three coarse modes, the paths taken and the nonzero counts."""
    out = {}
    for mode in (0, 1, 2):
        det.set_coarse_mode(mode)
        res = tuples(det.match(img, thr, class_ids))
        out[mode] = (res, det.last_coarse_path())
    det.set_coarse_mode(1)
    return out


@pytest.mark.parametrize("case", [0, 1, 2])
def test_fixture_banks_all_modes_equal_oracle(case):
    det = load_fixture_detector(case)
    img = crop16(scene_bgr(case))
    ref = oracle_tuples(oracle.match(img, [4, 8], 80, bank_of(det), nthreads=NTHREADS))
    out = run_modes(det, img, 80)
    for mode, (res, path) in out.items():
        say("case", case, "mode", mode, "path", path, "matches", len(res))
        assert res == ref, f"mode {mode}"
    assert out[0][1] == (False, -1)
    sparse2, nz = out[2][1]
    r, c = img.shape[0] // 2, img.shape[1] // 2
    assert nz > 0
    assert sparse2 == (nz <= r * c)  # the list holds up to one response in eight
    assert not out[1][1][0]  # auto mode keeps small coarse stages on the gather kernel


def test_synthetic_banks_sparse_and_dense_scenes():
    _, _, poly, _ = synth.scene_plain(0)
    det = sbm.Detector(128, [4, 8])
    for ang in range(0, 360, 4):
        det.add_template_pyramid("a", synth.bank_template(poly, float(ang), 1.0, 60, 128))
    for ang in range(0, 360, 9):
        det.add_template_pyramid("b", synth.bank_template(poly, float(ang), 0.8, 60, 300))
    scenes = [synth.scene_copies(1)[0], synth.scene_plain(2)[0], synth.scene_sized(3, 960, 1280, poly, n_shapes=30)]
    for img in scenes:
        ref = None
        for thr in (80, 65):
            out = run_modes(det, img, thr)
            say(img.shape, "threshold", thr, {m: (len(r), p) for m, (r, p) in out.items()})
            assert out[0][0] == out[1][0] == out[2][0]
            assert out[2][1][0]  # these scenes' coarse levels fit the list
            if thr == 80:
                ref = out[0][0]
        assert ref is not None
    # one against the oracle
    ref = oracle_tuples(oracle.match(scenes[0], [4, 8], 80, bank_of(det), nthreads=NTHREADS))
    det.set_coarse_mode(2)
    assert tuples(det.match(scenes[0], 80)) == ref
    det.set_coarse_mode(1)


def test_two_segments_per_template_and_class_selection():
    """1936x2592: the coarse level has 121 x 162 = 19,602 positions, more than one score map holds."""
    _, _, poly, _ = synth.scene_plain(0)
    det = sbm.Detector(128, [4, 8])
    for ang in range(0, 360, 10):
        det.add_template_pyramid("a", synth.bank_template(poly, float(ang), 1.0, 60, 128))
        det.add_template_pyramid("b", synth.bank_template(poly, float(ang) + 5, 1.1, 60, 200))
    img = synth.scene_sized(5, 1936, 2592, poly, n_shapes=40)
    for ids in (None, ["b"]):
        out = run_modes(det, img, 75, ids)
        say("1936x2592", ids, {m: (len(r), p) for m, (r, p) in out.items()})
        assert out[0][0] == out[2][0] == out[1][0]
        assert out[2][1][0]
        assert len(out[0][0]) > 0


def test_graph_replays_and_batches_take_the_same_path():
    _, _, poly, _ = synth.scene_plain(0)
    det = sbm.Detector(128, [4, 8])
    for ang in range(0, 360, 6):
        det.add_template_pyramid("a", synth.bank_template(poly, float(ang), 1.0, 60, 128))
    imgs = [synth.scene_copies(10 + k)[0] for k in range(3)] + [crop16(scene_bgr(2))[:, :, 0].copy()]
    det.set_coarse_mode(0)
    ref = [tuples(det.match(im, 75)) for im in imgs]
    det.set_coarse_mode(2)  # the list path wherever it fits (all but the last, textured, image)
    dimgs = [det.upload_image(im) for im in imgs]
    for rep in range(3):  # eager, captured, replayed
        for k, d in enumerate(dimgs):
            assert tuples(det.match_image(d, 75)) == ref[k], f"image {k} call {rep}"
    got = det.match_images(dimgs, 75, slots=3)
    for k in range(len(imgs)):
        assert tuples(got[k]) == ref[k]
    det.set_coarse_mode(1)


def test_auto_mode_takes_the_list_on_a_large_sparse_stage():
    """2,000 templates x 300 coarse features over a 1024x1024 scene: features x positions reach the auto
    mode's bar, and the coarse level is about 3% nonzero."""
    _, _, poly, _ = synth.scene_plain(0)
    det = sbm.Detector(128, [4, 8])
    for k in range(2000):
        det.add_template_pyramid("big", synth.bank_template(poly, k * 0.18, 0.9 + 0.1 * (k % 3), 60, 600))
    img = synth.scene_copies(1)[0]
    out = run_modes(det, img, 80)
    say({m: (len(r), p) for m, (r, p) in out.items()})
    assert out[0][0] == out[1][0] == out[2][0]
    assert out[1][1][0] and out[2][1][0] and not out[0][1][0]


@pytest.mark.parametrize("T", [[5], [2, 4, 8], [4, 8, 16]])
def test_other_pyramids_list_path_equals_oracle(T):
    """One level (the list feeds the final candidates directly) and three levels (two refinements after
    the row sort); masked, against the oracle."""
    _, templ, poly, _ = synth.scene_plain(0)
    det = sbm.Detector(64, T)
    big = T[-1] == 16  # the coarsest level must stay sparse under a 16 x 16 spread
    img = synth.scene_sized(21, 1024 if big else 480, 1280 if big else 640, poly, n_shapes=3, n_copies=2)
    rows, cols = img.shape
    yy, xx = np.mgrid[0:rows, 0:cols]
    mask = (((xx - cols / 2) ** 2 + (yy - rows / 2) ** 2) < (0.45 * rows) ** 2).astype(np.uint8) * 255
    base = det.addTemplate(templ, "poly", None, 64)
    assert base == 0
    for ang in range(10, 360, 10):
        det.addTemplate_rotate("poly", 0, float(ang), (80.0, 80.0))
    ref = oracle_tuples(oracle.match(img, T, 70, bank_of(det), mask=mask, nthreads=NTHREADS))
    for mode in (0, 2):
        det.set_coarse_mode(mode)
        got = tuples(det.match(img, 70, masks=mask))
        say("T", T, "mode", mode, "path", det.last_coarse_path(), "matches", len(got))
        assert got == ref, f"mode {mode}"
    assert det.last_coarse_path()[0]
    det.set_coarse_mode(1)


def test_candidate_overflow_grows_the_buffers_on_both_paths():
    """A low threshold (12) over 400 templates passes more coarse candidates than the initial buffers hold
    (65,536): the match grows them and runs again, eagerly and after graph replays, on both paths."""
    _, _, poly, _ = synth.scene_plain(0)
    det = sbm.Detector(128, [4, 8])
    for k in range(400):
        det.add_template_pyramid("p", synth.bank_template(poly, k * 0.9, 0.9 + 0.05 * (k % 5), 60, 256))
    img = synth.scene_copies(1)[0]
    res = {}
    for mode in (0, 2):
        d2 = sbm.Detector(128, [4, 8])
        for k in range(400):
            d2.add_template_pyramid("p", synth.bank_template(poly, k * 0.9, 0.9 + 0.05 * (k % 5), 60, 256))
        d2.set_coarse_mode(mode)
        dimg = d2.upload_image(img)
        outs = [tuples(d2.match_image(dimg, 12)) for _ in range(3)]  # eager (grows), captured, replayed
        cc = d2.last_timing()["coarse_candidates"]
        say("mode", mode, "coarse candidates", cc, "matches", len(outs[0]), "path", d2.last_coarse_path())
        assert cc > 65536
        assert outs[0] == outs[1] == outs[2]
        res[mode] = outs[0]
    assert res[0] == res[2]
    assert tuples(det.match(img, 12)) == res[0]


def test_odd_coarse_level_small_and_large_features():
    """1936x1936: the coarse level has 121 x 121 = 14,641 positions, odd and not a multiple of four, in one
    segment; templates of 16 to 1,024 coarse features.  The list path (four words per lane in the threshold
    scan, rows read from four-aligned slice starts, the sentinel after the list) equals the gather kernel,
    and the oracle at one threshold."""
    _, _, poly, _ = synth.scene_plain(0)
    det = sbm.Detector(128, [4, 8])
    for ang in range(0, 360, 15):
        det.add_template_pyramid("s", synth.bank_template(poly, float(ang), 1.0, 60, 32))
        det.add_template_pyramid("l", synth.bank_template(poly, float(ang) + 7, 1.1, 60, 2048))
    img = synth.scene_sized(7, 1936, 1936, poly, n_shapes=30)
    for thr in (80, 55):
        out = run_modes(det, img, thr)
        say("1936x1936 threshold", thr, {m: (len(r), p) for m, (r, p) in out.items()})
        assert out[0][0] == out[1][0] == out[2][0]
        assert out[2][1][0]
        assert len(out[0][0]) > 0
    ref = oracle_tuples(oracle.match(img, [4, 8], 80, bank_of(det), nthreads=NTHREADS))
    det.set_coarse_mode(2)
    assert tuples(det.match(img, 80)) == ref
    det.set_coarse_mode(1)


def test_autotune_searches_the_list_path_and_refinement_shapes():
    """On a workload whose coarse stage takes the nonzero-response list, the autotuner times the scatter's 17
    launch shapes (threads, segment, blocks per CU; items from the work counter), then three other splits of the
    pruned scatter on the best shape, and the refinement's 140, every one returning the same
    matches (it checks that itself), and keeps the fastest; each shape set by hand matches too."""
    _, _, poly, _ = synth.scene_plain(0)
    det = sbm.Detector(128, [4, 8])
    for k in range(2000):
        det.add_template_pyramid("big", synth.bank_template(poly, k * 0.18, 0.9 + 0.1 * (k % 3), 60, 600))
    img = synth.scene_copies(1)[0]
    dimg = det.upload_image(img)
    before = tuples(det.match_image(dimg, 80))
    assert det.last_coarse_path()[0]
    _, rows = det.autotune(dimg, 80, iters=1)
    lst = [r for r in rows if r["stage"] == "list"]
    ref = [r for r in rows if r["stage"] == "refine"]
    say("list shapes", len(lst), "refinement shapes", len(ref), "fastest list shape",
        min(lst, key=lambda r: r["us"]))
    assert len(lst) == 20 and len(ref) == 140 and not [r for r in rows if r["stage"] == "gather"]
    assert {r["list_order"] for r in lst} == {0} and {r["list_split"] for r in lst} == {100, 60, 50, 40}
    keys = ("list_threads", "list_segment", "list_blocks_per_cu", "list_order", "list_split")
    chosen = tuple(det.get_tuning_param(k) for k in keys)
    assert chosen in [tuple(r[k] for k in keys) for r in lst]
    assert tuples(det.match_image(dimg, 80)) == before
    for th, seg, bpc, order, split in ((512, 4096, 4, 1, 50), (1024, 8192, 1, 0, 100), (256, 4096, 4, 1, 40),
                                       (256, 16384, 1, 0, 60), (1024, 8192, 2, 1, 10)):
        det.set_tuning_param("list_blocks_per_cu", 1)  # every (threads, segment) fits one block per CU
        det.set_tuning_param("list_threads", th)
        det.set_tuning_param("list_segment", seg)
        det.set_tuning_param("list_blocks_per_cu", bpc)
        det.set_tuning_param("list_order", order)
        det.set_tuning_param("list_split", split)
        assert tuples(det.match_image(dimg, 80)) == before and det.last_coarse_path()[0]
    det.set_tuning_param("list_blocks_per_cu", 1)
    det.set_tuning_param("list_threads", 1024)
    with pytest.raises(sbm.SbmError):
        det.set_tuning_param("list_blocks_per_cu", 4)  # 1024 threads x 4 blocks exceed a CU
    det.set_tuning_param("list_segment", 16384)
    with pytest.raises(sbm.SbmError):
        det.set_tuning_param("list_blocks_per_cu", 2)  # two 16384-position maps (128 KB each) exceed its LDS
    with pytest.raises(sbm.SbmError):
        det.set_tuning_param("list_order", 2)
    with pytest.raises(sbm.SbmError):
        det.set_tuning_param("list_split", 5)
    with pytest.raises(sbm.SbmError):
        det.set_tuning_param("refine_waves", 5)


def test_pruned_scatter_every_split_equals_oracle():
    """The pruned scatter (list_split < 100): after that percent of a template's features the positions that can
    still pass are kept and the rest of the features gathered at them.  Every split returns the oracle's list,
    those below the point where the remaining features alone could pass (4 x remaining >= raw threshold) by
    scattering every feature."""
    _, _, poly, _ = synth.scene_plain(0)
    det = sbm.Detector(128, [4, 8])
    for k in range(120):
        det.add_template_pyramid("p", synth.bank_template(poly, k * 3.0, 0.9 + 0.1 * (k % 3), 60, 256))
    img = synth.scene_copies(1)[0]
    ref = oracle_tuples(oracle.match(img, [4, 8], 75, bank_of(det), nthreads=NTHREADS))
    assert len(ref) > 0
    det.set_coarse_mode(2)
    for split in (100, 60, 50, 40, 25, 10):
        det.set_tuning_param("list_split", split)
        got = tuples(det.match(img, 75))
        say("split", split, "matches", len(got), "entries read", det.last_coarse_entries())
        assert det.last_coarse_path()[0]
        assert got == ref, f"split {split}"
    det.set_tuning_param("list_split", 100)
    det.set_coarse_mode(1)


def test_pruned_scatter_overflowing_kept_positions_scatters_the_rest():
    """Vertical stripes every four pixels over the top quarter of a 1024^2 image give its coarse positions one
    orientation (the coarse level stays 10% nonzero, so the list holds it), and a template whose features all carry
    it keeps about 960 positions after half its features: the kept-position list (256 entries) overflows and the
    segment scatters its remaining features.  Lists equal the oracle's (708 matches) and the unpruned scatter's."""
    rows, cols = 1024, 1024
    img = np.zeros((rows, cols), np.uint8)
    img[:260, (np.arange(cols) // 4) % 2 == 1] = 200
    det = sbm.Detector(64, [4, 8])
    xs, ys = np.meshgrid(np.arange(8, 56, 6), np.arange(8, 56, 6))
    f0 = np.column_stack([xs.ravel(), ys.ravel(), np.zeros(xs.size)]).astype(np.int32)
    f1 = np.column_stack([xs.ravel() // 2, ys.ravel() // 2, np.zeros(xs.size)]).astype(np.int32)
    det.add_template_pyramid("stripes", [dict(width=56, height=56, tl_x=0, tl_y=0, features=f0),
                                         dict(width=28, height=28, tl_x=0, tl_y=0, features=f1)])
    ref = oracle_tuples(oracle.match(img, [4, 8], 90, bank_of(det), nthreads=NTHREADS))
    say("oracle matches", len(ref))
    assert len(ref) > 600  # more positions pass than the kept-position list holds
    det.set_coarse_mode(2)
    for split in (50, 100, 50):
        det.set_tuning_param("list_split", split)
        got = tuples(det.match(img, 90))
        assert det.last_coarse_path()[0]
        assert got == ref, f"split {split}"
    det.set_tuning_param("list_split", 100)
    det.set_coarse_mode(1)
