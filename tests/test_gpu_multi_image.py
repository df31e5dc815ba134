"""Several images in one launch of every stage (sbm_match_batch with images_per_launch > 1).

The preprocessing kernels take one grid slice per image, the coarse stage (image, tile, template) items, the
refinement the candidates of every image (the image in the template index's high bits); the nonzero-response
path decides per image.  Every image's list must equal the image matched alone, and the oracle, for every
launch size, slot count, tile shape and coarse mode, with masks, colour, other pyramids, 16 orientations, and
when the candidates overflow the captured buffers.
"""
import os

import numpy as np
import pytest

import oracle
import shape_based_matching_amd as sbm
from helpers import bank_of, oracle_tuples, tuples
from shape_based_matching_amd import synth

pytestmark = pytest.mark.gpu
NTHREADS = max(1, min(16, os.cpu_count() or 1))


def say(*a):
    print(*a, flush=True)


def poly_detector(nf=128, T=(4, 8), n=60, step=6, feats=128, bins=8):
    _, _, poly, _ = synth.scene_plain(0)
    det = sbm.Detector16(nf, list(T)) if bins == 16 else sbm.Detector(nf, list(T))
    for k in range(n):
        det.add_template_pyramid("a", synth.bank_template(poly, float(k * step), 1.0, 60, feats, bins=bins))
    return det, poly


def check_batches(det, dimgs, thr, alone, per_launch=(1, 2, 3, 8, 16), slots=(1, 2, 4)):
    for per in per_launch:
        det.set_images_per_launch(per)
        for sl in slots:
            for rep in range(2):  # captured, replayed
                got = det.match_images(dimgs, thr, slots=sl)
                for k in range(len(dimgs)):
                    assert tuples(got[k]) == alone[k], f"images per launch {per} slots {sl} pass {rep} image {k}"
    det.set_images_per_launch(0)
    got = det.match_images(dimgs, thr, slots=2)
    for k in range(len(dimgs)):
        assert tuples(got[k]) == alone[k], f"automatic launch size, image {k}"


def test_one_size_many_images_all_launch_sizes():
    det, poly = poly_detector()
    imgs = [synth.scene_copies(40 + k)[0] for k in range(9)]
    dimgs = [det.upload_image(im) for im in imgs]
    alone = [tuples(det.match_image(d, 75)) for d in dimgs]
    say("matches alone", [len(a) for a in alone])
    assert sum(len(a) for a in alone) > 0
    check_batches(det, dimgs, 75, alone)
    ref = oracle_tuples(oracle.match(imgs[3], [4, 8], 75, bank_of(det), nthreads=NTHREADS))
    assert alone[3] == ref


@pytest.mark.parametrize("tuning", [(16, 4, 8, 2, 1), (32, 4, 4, 3, 4), (8, 2, 8, 0, 1), (4, 1, 16, 1, 2)])
def test_tile_shapes_in_multi_image_launches(tuning):
    det, poly = poly_detector(n=40, step=9)
    det.set_tuning(*tuning)
    imgs = [synth.scene_copies(60 + k)[0] for k in range(5)]
    dimgs = [det.upload_image(im) for im in imgs]
    alone = [tuples(det.match_image(d, 70)) for d in dimgs]
    check_batches(det, dimgs, 70, alone, per_launch=(5, 2), slots=(1, 3))


def test_timed_launches_run_eagerly_and_report_their_kernels():
    """With per-kernel timing a multi-image launch runs its stages eagerly: the lists equal the graph launches'
    and the image alone, and the kernel times and counters read afterwards are that launch's (every kernel
    covering all of its images)."""
    det, poly = poly_detector(n=360, step=1, feats=256)
    imgs = [synth.scene_copies(70 + k)[0] for k in range(8)]
    dimgs = [det.upload_image(im) for im in imgs]
    alone = [tuples(det.match_image(d, 80)) for d in dimgs]
    det.set_images_per_launch(8)
    graph = [tuples(r) for r in det.match_images(dimgs, 80, slots=1)]
    assert graph == alone
    det.set_timing(True)
    for rep in range(2):
        timed = [tuples(r) for r in det.match_images(dimgs, 80, slots=1)]
        assert timed == alone, f"timed pass {rep}"
        names = [n for n, _ in det.last_kernel_times()]
        say("timed launch kernels", det.last_kernel_times())
        assert {"gradient_l0", "coarse_batched", "refine_l0"} <= set(names)
        assert all(ms > 0 for _, ms in det.last_kernel_times())
        assert det.last_refine_features() > 0
    det.set_timing(False)
    assert [tuples(r) for r in det.match_images(dimgs, 80, slots=2)] == alone
    det.set_images_per_launch(0)


def test_sparse_and_dense_images_in_one_launch():
    """Coarse mode 2: the list path wherever the list fits.  Textured images (noise) do not fit and take the
    gather kernel in the same launches as sparse scenes that take the list."""
    det, poly = poly_detector(n=90, step=4)
    g = np.random.default_rng(5)
    imgs = []
    for k in range(6):
        if k % 3 == 1:
            imgs.append(g.integers(0, 256, (1024, 1024), dtype=np.uint8))  # every pixel an edge
        else:
            imgs.append(synth.scene_copies(80 + k)[0])
    dimgs = [det.upload_image(im) for im in imgs]
    det.set_coarse_mode(0)
    ref = [tuples(det.match_image(d, 70)) for d in dimgs]
    det.set_coarse_mode(2)
    paths = []
    for d in dimgs:
        det.match_image(d, 70)
        paths.append(det.last_coarse_path()[0])
    say("list path per image", paths)
    assert any(paths) and not all(paths)
    check_batches(det, dimgs, 70, ref, per_launch=(6, 4, 2), slots=(1, 2))
    det.set_coarse_mode(1)
    check_batches(det, dimgs, 70, ref, per_launch=(6,), slots=(1,))


def test_masks_colour_and_mixed_sizes():
    det, poly = poly_detector(n=60, step=6)
    imgs, masks = [], []
    for k in range(8):
        rows, cols = ((480, 640), (960, 1280))[k % 2]
        im = synth.scene_sized(200 + k, rows, cols, poly, n_shapes=15)
        m = np.zeros((rows, cols), np.uint8)
        m[20 + 4 * k:-30, 40:-40 - 8 * k] = 255
        imgs.append(im)
        masks.append(m)
    dimgs = [det.upload_image(im, mask=m) for im, m in zip(imgs, masks)]
    alone = [tuples(det.match(im, 70, masks=m)) for im, m in zip(imgs, masks)]
    check_batches(det, dimgs, 70, alone, per_launch=(4, 3), slots=(1, 2))
    ref = oracle_tuples(oracle.match(imgs[5], [4, 8], 70, bank_of(det), mask=masks[5], nthreads=NTHREADS))
    assert alone[5] == ref
    # colour images, no mask, together with the masked gray ones in one call (three size buckets)
    bgr = [np.ascontiguousarray(np.dstack([im, im[:, ::-1], 255 - im])) for im in imgs[:4]]
    dbgr = [det.upload_image(b) for b in bgr]
    alone_bgr = [tuples(det.match(b, 70)) for b in bgr]
    det.set_images_per_launch(2)
    got = det.match_images(dimgs + dbgr, 70, slots=3)
    for k in range(len(dimgs)):
        assert tuples(got[k]) == alone[k]
    for k in range(len(dbgr)):
        assert tuples(got[len(dimgs) + k]) == alone_bgr[k]
    det.set_images_per_launch(0)


@pytest.mark.parametrize("T", [[5, 8], [2, 4, 8]])
def test_other_pyramids_in_multi_image_launches(T):
    """T = 5 (the per-row spread kernel) and three levels (two refinements, a 2x2 spread)."""
    _, templ, poly, _ = synth.scene_plain(0)
    det = sbm.Detector(64, T)
    assert det.addTemplate(templ, "poly", None, 64) == 0
    for ang in range(10, 360, 10):
        det.addTemplate_rotate("poly", 0, float(ang), (80.0, 80.0))
    rows = 400 if T[0] == 5 else 480  # every level divisible by its T
    imgs = [synth.scene_sized(300 + k, rows, 640, poly, n_shapes=8) for k in range(4)]
    dimgs = [det.upload_image(im) for im in imgs]
    alone = [tuples(det.match_image(d, 65)) for d in dimgs]
    say("T", T, "matches alone", [len(a) for a in alone])
    check_batches(det, dimgs, 65, alone, per_launch=(4, 3), slots=(1, 2))
    ref = oracle_tuples(oracle.match(imgs[1], T, 65, bank_of(det), nthreads=NTHREADS))
    assert alone[1] == ref


def test_16_orientations_in_multi_image_launches():
    det, poly = poly_detector(nf=63, T=(4, 8), n=36, step=10, feats=128, bins=16)
    imgs = []
    for k in range(4):
        g = synth.scene_copies(400 + k)[0]
        imgs.append(np.ascontiguousarray(np.dstack([g, np.roll(g, 1, axis=1), g])))
    dimgs = [det.upload_image(im) for im in imgs]
    alone = [tuples(det.match_image(d, 55)) for d in dimgs]
    say("16 orientations: matches alone", [len(a) for a in alone])
    check_batches(det, dimgs, 55, alone, per_launch=(4, 2), slots=(1, 2))
    ref = oracle_tuples(oracle.match16(imgs[2], [4, 8], 55, bank_of(det), nthreads=NTHREADS))
    assert alone[2] == ref


def test_candidate_overflow_in_a_multi_image_launch():
    """Threshold 12 over 400 templates: each image passes tens of thousands of coarse candidates, together more
    than the captured buffers hold; the launch is redone image by image with grown buffers, and the next
    launch captures again."""
    _, _, poly, _ = synth.scene_plain(0)
    det = sbm.Detector(128, [4, 8])
    for k in range(400):
        det.add_template_pyramid("p", synth.bank_template(poly, k * 0.9, 0.9 + 0.05 * (k % 5), 60, 256))
    imgs = [synth.scene_copies(500 + k)[0] for k in range(3)]
    dimgs = [det.upload_image(im) for im in imgs]
    alone = [tuples(det.match_image(d, 12)) for d in dimgs]
    say("coarse candidates of the last image alone", det.last_timing()["coarse_candidates"])
    fresh = sbm.Detector(128, [4, 8])
    for k in range(400):
        fresh.add_template_pyramid("p", synth.bank_template(poly, k * 0.9, 0.9 + 0.05 * (k % 5), 60, 256))
    fd = [fresh.upload_image(im) for im in imgs]
    fresh.set_images_per_launch(3)
    for rep in range(3):
        got = fresh.match_images(fd, 12, slots=1)
        for k in range(3):
            assert tuples(got[k]) == alone[k], f"pass {rep} image {k}"


def test_handle_and_array_entry_points_agree():
    """sbm_match_batch (one sbm_matches per image) and sbm_match_batch_arrays (the lists back to back in the
    caller's arrays, a cap below the total reported through the counts) return the same lists."""
    import ctypes

    from shape_based_matching_amd import _lib
    det, poly = poly_detector(n=30, step=12)
    imgs = [synth.scene_copies(600 + k)[0] for k in range(5)]
    dimgs = [det.upload_image(im) for im in imgs]
    alone = [tuples(det.match_image(d, 70)) for d in dimgs]
    det.set_images_per_launch(3)
    n = len(dimgs)
    arr = (ctypes.c_void_p * n)(*[d._h.value for d in dimgs])
    outs = (ctypes.c_void_p * n)()
    ids = (ctypes.c_char_p * 1)()
    _lib.check(_lib.lib().sbm_match_batch(det._h, arr, n, 70.0, ids, 0, 2, outs))
    for k in range(n):
        xyct, sim, names = det._collect(ctypes.c_void_p(outs[k]))
        got = [(int(r[0]), int(r[1]), np.float32(s), names[r[2]], int(r[3])) for r, s in zip(xyct, sim)]
        assert got == alone[k], k
    total = sum(len(a) for a in alone)
    counts = np.zeros(n, np.int32)
    cap = max(1, total // 2)
    xyct = np.zeros((cap, 4), np.int32)
    sim = np.zeros(cap, np.float32)
    _lib.check(_lib.lib().sbm_match_batch_arrays(det._h, arr, n, 70.0, ids, 0, 2, xyct.ctypes.data, sim.ctypes.data,
                                                 cap, counts.ctypes.data))
    assert list(counts) == [len(a) for a in alone]
    flat = [m for a in alone for m in a][:cap]
    names = det.classIds()
    assert [(int(r[0]), int(r[1]), np.float32(s), names[r[2]], int(r[3])) for r, s in zip(xyct, sim)] == flat


def test_host_image_batches():
    """sbm_match_batch_host_arrays: host images (row-padded ones among them), masks, colour and two sizes go
    through pinned staging, one DMA per launch; every list equals the image matched alone."""
    det, poly = poly_detector(n=40, step=9)
    imgs, masks = [], []
    for k in range(7):
        rows, cols = ((480, 640), (960, 1280))[k % 2]
        im = synth.scene_sized(700 + k, rows, cols, poly, n_shapes=12)
        if k == 3:
            padded = np.zeros((rows, cols + 24), np.uint8)
            padded[:, :cols] = im
            im = padded[:, :cols]  # a strided view: row stride cols + 24
        imgs.append(im)
        m = np.zeros((rows, cols), np.uint8)
        m[16:-16, 32:-32 - 8 * k] = 255
        masks.append(m)
    alone = [tuples(det.match(im, 70)) for im in imgs]
    alone_m = [tuples(det.match(im, 70, masks=m)) for im, m in zip(imgs, masks)]
    for per in (1, 3, 4):
        det.set_images_per_launch(per)
        for sl in (1, 2):
            got = det.match_host_images_arrays(imgs, 70, slots=sl)
            names = det.classIds()
            for k in range(len(imgs)):
                xyct, sim, _ = got[k]
                assert [(int(r[0]), int(r[1]), np.float32(s), names[r[2]], int(r[3])) for r, s in zip(xyct, sim)] \
                    == alone[k], (per, sl, k)
            got = det.match_host_images_arrays(imgs, 70, masks=masks, slots=sl)
            for k in range(len(imgs)):
                xyct, sim, _ = got[k]
                assert [(int(r[0]), int(r[1]), np.float32(s), names[r[2]], int(r[3])) for r, s in zip(xyct, sim)] \
                    == alone_m[k], ("mask", per, sl, k)
    det.set_images_per_launch(0)


def test_autotune_times_multi_image_launches():
    """With images_per_launch 8 the autotuner times launches of eight copies of the image, which on the 1024^2 x
    360-template x 256-feature workload go down the nonzero-response path (8 x 1.9e8 features x positions at the
    coarse level), while the image alone takes the gather kernel: it searches the list shapes on the launches and
    the gather kernel's on the same launches; if the gather kernel is faster there the workload keeps coarse mode
    0, else the gather shapes are searched again on single images.  Then the refinement's shapes; every launch of
    the stream still equals each image matched alone."""
    _, _, poly, _ = synth.scene_plain(0)
    det = sbm.Detector(128, [4, 8])
    for k in range(360):
        det.add_template_pyramid("poly", synth.bank_template(poly, float(k), 1.0, 60, 256))
    imgs = [synth.scene_copies(60 + k)[0] for k in range(8)]
    dimgs = [det.upload_image(im) for im in imgs]
    alone = [tuples(det.match_image(d, 80)) for d in dimgs]
    assert not det.last_coarse_path()[0] and sum(len(a) for a in alone) > 0
    det.set_images_per_launch(8)
    _, rows = det.autotune(dimgs[0], 80, iters=1)
    stages = {st: len([r for r in rows if r["stage"] == st]) for st in ("gather", "list", "refine")}
    say("stages timed", stages, "fastest launch", min(rows, key=lambda r: r["us"]))
    # gather rows: one probe shape on the launches, then the 240 shapes on them when the probe comes near the
    # list, then the 240 on single images when the list stays
    assert stages["list"] == 20 and stages["refine"] == 140 and stages["gather"] in (241, 481)
    if det.get_coarse_mode() == 0:
        assert stages["gather"] == 241
    for rep in range(2):
        got = det.match_images(dimgs, 80, slots=2)
        assert [tuples(g) for g in got] == alone, f"pass {rep}"
    det.set_images_per_launch(4)  # four to a launch stay on the gather kernel
    _, rows = det.autotune(dimgs[0], 80, iters=1)
    assert len([r for r in rows if r["stage"] == "gather"]) == 240
    assert [tuples(g) for g in det.match_images(dimgs, 80, slots=2)] == alone


def test_autotune_on_distinct_scenes():
    """sbm_autotune_images: the multi-image launches the autotuner times take the given scenes in turn (each
    checked against that scene matched alone) instead of copies of the first; the stream then still equals every
    scene matched alone and the oracle.  Scenes of another size, or masked, are refused before any timing."""
    det, poly = poly_detector(n=60)
    imgs = [synth.scene_copies(70 + k)[0] for k in range(4)]
    dimgs = [det.upload_image(im) for im in imgs]
    alone = [tuples(det.match_image(d, 80)) for d in dimgs]
    assert alone[1] == oracle_tuples(oracle.match(imgs[1], [4, 8], 80, bank_of(det), nthreads=NTHREADS))
    det.set_images_per_launch(4)
    _, rows = det.autotune(dimgs[0], 80, iters=1, batch=dimgs[1:])
    say("rows", len(rows), "fastest", min(rows, key=lambda r: r["us"]))
    assert len([r for r in rows if r["stage"] == "refine"]) == 140
    for rep in range(2):
        assert [tuples(g) for g in det.match_images(dimgs, 80, slots=2)] == alone, f"pass {rep}"
    small = det.upload_image(imgs[0][:512, :512])
    masked = det.upload_image(imgs[2], mask=np.ones(imgs[2].shape[:2], np.uint8))
    for bad in (small, masked):
        with pytest.raises(sbm.SbmError):
            det.autotune(dimgs[0], 80, iters=1, batch=[dimgs[1], bad])
    det.set_images_per_launch(0)
