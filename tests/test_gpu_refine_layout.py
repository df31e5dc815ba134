"""The refinement's layouts: the banded copy of each refined level with 8-byte loads (refine_layout 1) and with
16-byte column loads (3), 16-byte loads over the linear memories (2), and the nibble-banded copy with 16-byte
slot loads (4).

The copy (k_band) holds, for each plane-phase, band of 16 grid rows and column 0 .. W + 15, the 32 linear
bytes of that column's rows 16 b .. 16 b + 31, reads past a row's end, past the plane's end and past the
linear block (zeros) included, so every 16 x 16 window is one run of 512 bytes.  The nibble-banded copy
(k_band_nib) holds the same rows two per byte (row 2k in the low nibble of slot byte k), 16 bytes per slot.
Each must equal its construction from the linear memories byte for byte, and every match list with the banded refinement must
equal the oracle's and the linear refinement's: gray and colour fixture scenes, masks, pyramids {4, 8},
{5, 8}, {2, 4, 8} (two refined levels) and {4, 2} (a refined level whose W is odd: the byte-load path of the
copy), 16 orientations, 8191-feature templates, multi-image launches and graph replays.
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


def crop(img, m):
    r, c = img.shape[:2]
    return np.ascontiguousarray(img[: r - r % m, : c - c % m])


def band_from_linear(lm, planes, T, W, H):
    """The banded copy built from the linear memories [planes*T*T][H*W] (zeros past the block)."""
    nb, wp = (H + 15) // 16, W + 16
    ext = np.concatenate([lm, np.zeros(40 * W + 64, np.uint8)])
    pp = np.arange(planes * T * T)[:, None, None, None]
    b = np.arange(nb)[None, :, None, None]
    c = np.arange(wp)[None, None, :, None]
    s = np.arange(32)[None, None, None, :]
    band = ext[pp * H * W + (16 * b + s) * W + c]
    return np.concatenate([band.reshape(-1), np.zeros(1024, np.uint8)])


def nib_band_from_linear(lm, planes, T, W, H):
    """The nibble-banded copy: the banded copy's 32-byte slots with rows 2k and 2k + 1 in byte k (low, high)."""
    band = band_from_linear(lm, planes, T, W, H)[:-1024].reshape(-1, 32)
    assert band.max(initial=0) < 16  # responses are 0..4
    nib = band[:, 0::2] | (band[:, 1::2] << 4)
    return np.concatenate([nib.reshape(-1), np.zeros(1024, np.uint8)])


LAYOUTS = (0, 1, 2, 3, 4)


def all_layouts(det, fn):
    out = []
    for layout in LAYOUTS:
        det.set_tuning_param("refine_layout", layout)
        out.append(fn())
    det.set_tuning_param("refine_layout", 0)
    return out


@pytest.mark.parametrize("layout", [1, 4])
@pytest.mark.parametrize("T,shape,planes", [([4, 8], (480, 640), 8), ([2, 4, 8], (480, 640), 8),
                                            ([5, 8], (400, 640), 8), ([4, 2], (240, 324), 8),
                                            ([5, 8], (400, 640), 16)])
def test_banded_copy_equals_its_construction(T, shape, planes, layout):
    _, _, poly, _ = synth.scene_plain(0)
    img = synth.scene_sized(21, shape[0], shape[1], poly, n_shapes=20)
    if planes == 16:
        img = np.ascontiguousarray(np.dstack([img, img[:, ::-1], 255 - img]))
    det = sbm.Detector16(T) if planes == 16 else sbm.Detector(63, T)
    det.set_tuning_param("refine_layout", layout)
    build = nib_band_from_linear if layout == 4 else band_from_linear
    mask = np.zeros(shape, np.uint8)
    mask[20:-30, 40:-20] = 255
    for m in (None, mask):
        det.prepare_image(img, m)
        for lvl in range(len(T) - 1):
            r, c = det.level_shape(lvl)
            Tl = T[lvl]
            W, H = c // Tl, r // Tl
            ref = build(det.read_linear_memories(lvl), planes, Tl, W, H)
            got = det.read_band(lvl)
            say("layout", layout, "T", T, "level", lvl, "W", W, "H", H, "bytes", got.size, "mask", m is not None)
            assert got.size == ref.size and np.array_equal(got, ref), (T, lvl)
        with pytest.raises(sbm.SbmError):
            det.read_band(len(T) - 1)  # the coarse level has no banded copy


@pytest.mark.parametrize("case", [0, 1, 2])
def test_fixture_scenes_all_layouts_equal_oracle(case):
    det = load_fixture_detector(case)
    bgr = crop(scene_bgr(case), 16)
    for img in (to_gray(bgr), bgr):
        ref = oracle_tuples(oracle.match(img, [4, 8], 70, bank_of(det), nthreads=NTHREADS))
        got = all_layouts(det, lambda: tuples(det.match(img, 70)))
        say("case", case, "channels", img.ndim, "matches", len(ref))
        assert got == [ref] * len(LAYOUTS)


@pytest.mark.parametrize("T", [[2, 4, 8], [5, 8], [4, 2]])
def test_other_pyramids_with_masks_equal_oracle(T):
    _, templ, poly, _ = synth.scene_plain(0)
    rows, cols = (400, 640) if 5 in T else ((240, 324) if T == [4, 2] else (480, 640))
    img = synth.scene_sized(33, rows, cols, poly, n_shapes=25)
    det = sbm.Detector(64, T)
    assert det.addTemplate(templ, "poly", None, 64) == 0
    for ang in range(10, 360, 10):
        det.addTemplate_rotate("poly", 0, float(ang), (80.0, 80.0))
    mask = np.zeros((rows, cols), np.uint8)
    mask[16:-40, 24:-8] = 255
    for m in (None, mask):
        ref = oracle_tuples(oracle.match(img, T, 60, bank_of(det), mask=m, nthreads=NTHREADS))
        got = all_layouts(det, lambda: tuples(det.match(img, 60, masks=m)))
        say("T", T, "mask", m is not None, "matches", len(ref))
        assert got == [ref] * len(LAYOUTS)


def test_16_orientations_all_layouts_equal_oracle():
    _, _, poly, _ = synth.scene_plain(0)
    img = synth.scene_sized(41, 400, 640, poly, n_shapes=25)
    bgr = np.ascontiguousarray(np.dstack([img, img[:, ::-1], 255 - img]))
    det = sbm.Detector16([5, 8])
    for ang in range(0, 360, 12):
        det.add_template_pyramid("p", synth.bank_template(poly, float(ang), 1.0, 50, 96, bins=16))
    ref = oracle_tuples(oracle.match16(bgr, [5, 8], 50, bank_of(det), nthreads=NTHREADS))
    got = all_layouts(det, lambda: tuples(det.match(bgr, 50)))
    say("16 orientations: matches", len(ref))
    assert got == [ref] * len(LAYOUTS)


def test_limit_templates_batches_and_replays():
    """8191-feature levels, eight images per launch and repeated (replayed) calls, on every refinement layout."""
    g = np.random.default_rng(5)
    _, _, poly, _ = synth.scene_plain(0)
    det = sbm.Detector(128, [4, 8])
    bank = {}
    f = np.column_stack([np.full(8191, 20), np.full(8191, 20), np.arange(8191) % 8]).astype(np.int32)
    r = np.column_stack([g.integers(0, 60, 8191), g.integers(0, 60, 8191), g.integers(0, 8, 8191)]).astype(np.int32)
    for pyr in ([dict(width=40, height=40, tl_x=0, tl_y=0, features=f),
                 dict(width=20, height=20, tl_x=0, tl_y=0, features=f)],
                [dict(width=60, height=60, tl_x=0, tl_y=0, features=r),
                 dict(width=30, height=30, tl_x=0, tl_y=0, features=r[:4000] // [2, 2, 1])]):
        det.add_template_pyramid("lim", pyr)
        bank.setdefault("lim", []).append(pyr)
    for ang in range(0, 360, 6):
        det.add_template_pyramid("p", synth.bank_template(poly, float(ang), 1.0, 60, 256))
    imgs = [synth.scene_copies(70 + k)[0] for k in range(8)]
    dimgs = [det.upload_image(im) for im in imgs]
    lin = [tuples(det.match_image(d, 30)) for d in dimgs]
    ref = oracle_tuples(oracle.match(imgs[0], [4, 8], 30, bank_of(det), nthreads=NTHREADS))
    assert lin[0] == ref and len(ref) > 0
    for layout in (1, 2, 3, 4):
        det.set_tuning_param("refine_layout", layout)
        for rep in range(2):  # captured, then replayed
            assert [tuples(det.match_image(d, 30)) for d in dimgs] == lin, f"layout {layout} single, pass {rep}"
        for per in (8, 3):
            det.set_images_per_launch(per)
            for rep in range(2):
                got = det.match_images(dimgs, 30, slots=2)
                assert [tuples(x) for x in got] == lin, f"layout {layout}, {per} per launch, pass {rep}"
        det.set_images_per_launch(0)
    det.set_tuning_param("refine_layout", 0)


@pytest.mark.parametrize("layout", [1, 3, 4])
def test_banded_copy_after_a_larger_image(layout):
    """A pipeline's banded buffer that held a larger image's copy: the zeroed block of the smaller copy lies
    inside the larger one's slots and is rewritten on every pass, so features outside the small image still
    read zeros.  Each image's list equals its 8-byte linear refinement's and the oracle's."""
    _, _, poly, _ = synth.scene_plain(0)
    det = sbm.Detector(128, [4, 8])
    for ang in range(0, 360, 5):
        det.add_template_pyramid("p", synth.bank_template(poly, float(ang), 1.0, 60, 256))
    big = synth.scene_sized(51, 960, 1280, poly, n_shapes=30)
    small = synth.scene_sized(52, 480, 640, poly, n_shapes=20)
    ref_small = oracle_tuples(oracle.match(small, [4, 8], 60, bank_of(det), nthreads=NTHREADS))
    lin = [tuples(det.match(im, 60)) for im in (big, small)]
    assert lin[1] == ref_small and len(ref_small) > 0
    det.set_tuning_param("refine_layout", layout)
    for rep in range(2):
        got = [tuples(det.match(im, 60)) for im in (big, small, big, small)]
        assert got == lin + lin, f"pass {rep}"
    dimgs = [det.upload_image(im) for im in (big, small, small, big)]
    det.set_images_per_launch(2)
    got = det.match_images(dimgs, 60, slots=1)  # one pipeline: the 960 x 1280 launch, then the 480 x 640 one
    assert [tuples(x) for x in got] == [lin[0], lin[1], lin[1], lin[0]]
    det.set_images_per_launch(0)
    det.set_tuning_param("refine_layout", 0)
